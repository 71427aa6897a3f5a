// Continuous-batching search server (see server.h).
#include "server.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "kernels.h"

namespace nsparse_amd {

namespace {

char* pinned(size_t bytes) {
    void* p = nullptr;
    hip_check(hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<char*>(p);
}

// Arrays inside a slot's blocks start on 256 bytes.
size_t up256(size_t b) { return (b + 255) & ~size_t{255}; }

// Zero bytes after the term ids and after every other input array (vector
// loads may run past an array's end), and the offset of the first output array.
constexpr size_t kIxPad = 64, kPad = 16, kOutHead = 256;

// Input block bytes for n queries with nnz non-zeros in all; the byte offsets
// of the k, selector, term-id and value arrays go to ks, sel, ix and v when
// they are given (the query offsets start the block).
size_t input_layout(int64_t n, int64_t nnz, size_t* ks, size_t* sel, size_t* ix, size_t* v) {
    const size_t o_ks = up256((n + 1) * 4 + kPad);
    const size_t o_sel = o_ks + up256(n * 4 + kPad);
    const size_t o_ix = o_sel + up256(n * 4 + kPad);
    const size_t o_v = o_ix + up256(nnz * 2 + kIxPad);
    if (ks != nullptr) {
        *ks = o_ks;
        *sel = o_sel;
        *ix = o_ix;
        *v = o_v;
    }
    return o_v + nnz * 4 + kPad;
}

}  // namespace

Server::Server(SeismicIndex& index, const ServerConfig& cfg) : index_(index), cfg_(cfg) {
    if (cfg_.max_batch <= 0 || cfg_.n_streams <= 0 || cfg_.k_max <= 0 || cfg_.scratch_budget <= 0)
        throw std::invalid_argument("server sizes must be positive");
    if (cfg_.n_streams > 16) throw std::invalid_argument("at most 16 server streams");
    if (cfg_.req.cut <= 0) throw std::invalid_argument("value must be positive");
    if (request_bytes(kDeviceQueryMax) > cfg_.scratch_budget)
        throw std::invalid_argument("scratch budget below one full-length request");
    if (const char* e = std::getenv("NSPARSE_SERVE_SPREAD_LOAD")) spread_load_ = std::atoll(e);
    require_device();
    (void)index_.view();  // index resident in HBM before the first batch
    epoch_ = std::chrono::steady_clock::now();
    slots_.resize(cfg_.n_streams);
    const size_t B = static_cast<size_t>(cfg_.max_batch), Q = B * kDeviceQueryMax, O = B * cfg_.k_max;
    for (Slot& s : slots_) {
        // default priority: a stream of any other priority is never used here
        hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate");
        const size_t in_max = input_layout(static_cast<int64_t>(B), static_cast<int64_t>(Q), nullptr, nullptr,
                                           nullptr, nullptr);
        const size_t out_max = kOutHead + up256(O * 4) + O * 4;
        s.h_in = pinned(in_max);
        s.h_out = pinned(out_max);
        s.d_in.allocate(in_max);
        s.d_out.allocate(out_max);
    }
    worker_ = std::thread([this] { run(); });
}

Server::~Server() {
    stop_ = true;
    cv_submit_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (Slot& s : slots_) {
        if (s.stream != nullptr) (void)hipStreamSynchronize(s.stream);
        (void)hipHostFree(s.h_in);
        (void)hipHostFree(s.h_out);
        if (s.done != nullptr) (void)hipEventDestroy(s.done);
        if (s.stream != nullptr) (void)hipStreamDestroy(s.stream);
    }
}

int64_t Server::now_ns() const {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - epoch_).count();
}

int64_t Server::request_bytes(int32_t nnz) const {
    return static_cast<int64_t>(nnz) * 6 + 12 + static_cast<int64_t>(cfg_.k_max) * 8;
}

int32_t Server::add_bitmap(const uint32_t* words, int64_t n_words) {
    if (words == nullptr || n_words <= 0) throw std::invalid_argument("empty selector bitmap");
    const int64_t need = (static_cast<int64_t>(index_.view().n_docs) + 31) / 32;
    if (n_words < need) throw std::invalid_argument("selector bitmap shorter than the index");
    std::lock_guard<std::mutex> lk(mu_);
    if (!pending_.empty() || inflight_bytes_ != 0)
        throw std::runtime_error("register selector bitmaps before submitting requests");
    DeviceBuffer b;
    b.upload(words, static_cast<size_t>(need) * 4);
    bitmap_ptrs_.push_back(b.get<uint32_t>());
    bitmaps_.push_back(std::move(b));
    d_bitmap_table_.upload(bitmap_ptrs_);
    return static_cast<int32_t>(bitmaps_.size() - 1);
}

int64_t Server::submit(int32_t nnz, const uint16_t* indices, const float* values, int32_t k, int32_t bitmap_id,
                       int64_t arrival_ns) {
    if (nnz < 0 || nnz > kDeviceQueryMax)
        throw std::invalid_argument("query has more non-zeros than the device path supports (512)");
    if (nnz > 0 && (indices == nullptr || values == nullptr)) throw std::invalid_argument("unexpected nullptr");
    if (k <= 0 || k > cfg_.k_max) throw std::invalid_argument("k outside [1, k_max] of the server");
    const int32_t dim = index_.view().dim;
    for (int32_t i = 0; i < nnz; ++i)
        if (indices[i] >= dim) throw std::invalid_argument("query term id out of range");
    Request r;
    r.indices.assign(indices, indices + nnz);
    r.values.assign(values, values + nnz);
    r.k = k;
    r.arrival_ns = arrival_ns < 0 ? now_ns() : arrival_ns;
    std::lock_guard<std::mutex> lk(mu_);
    if (bitmap_id < -1 || bitmap_id >= static_cast<int32_t>(bitmaps_.size()))
        throw std::invalid_argument("unknown selector bitmap id");
    if (!error_.empty()) throw std::runtime_error(error_);
    r.bitmap_id = bitmap_id;
    requests_.push_back(std::move(r));
    const int64_t ticket = first_ticket_ + static_cast<int64_t>(requests_.size()) - 1;
    pending_.push_back(ticket);
    cv_submit_.notify_one();
    return ticket;
}

double Server::wait(int64_t ticket, float* distances, int32_t* labels) {
    std::unique_lock<std::mutex> lk(mu_);
    if (ticket < first_ticket_ || ticket >= first_ticket_ + static_cast<int64_t>(requests_.size()))
        throw std::invalid_argument("unknown or already collected ticket");
    cv_done_.wait(lk, [&] { return requests_[ticket - first_ticket_].done_ns >= 0 || !error_.empty(); });
    Request& r = requests_[ticket - first_ticket_];
    if (r.done_ns < 0) throw std::runtime_error(error_);
    if (distances != nullptr) std::copy(r.dist.begin(), r.dist.end(), distances);
    if (labels != nullptr) std::copy(r.lab.begin(), r.lab.end(), labels);
    const double us = static_cast<double>(r.done_ns - r.arrival_ns) / 1e3;
    r.collected = true;
    r.indices.clear();
    r.values.clear();
    while (!requests_.empty() && requests_.front().collected) {
        requests_.pop_front();
        ++first_ticket_;
    }
    return us;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

void Server::run() {
    try {
        while (!stop_) {
            bool progress = reap();
            progress = admit() || progress;
            if (!progress) {
                std::unique_lock<std::mutex> lk(mu_);
                cv_submit_.wait_for(lk, std::chrono::microseconds(20));
            }
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu_);
        error_ = std::string("server dispatcher: ") + e.what();
        cv_done_.notify_all();
    }
}

// Finished batches: results out of pinned memory into their requests.
bool Server::reap() {
    bool any = false;
    for (Slot& s : slots_) {
        if (!s.busy) continue;
        const hipError_t e = hipEventQuery(s.done);
        if (e == hipErrorNotReady) continue;
        hip_check(e, "server batch");
        const int32_t err = *reinterpret_cast<const int32_t*>(s.h_out);
        if (err != 0) throw std::runtime_error("server batch: " + device_error_message(err));
        const int64_t t_done = now_ns();
        const float* h_dist = reinterpret_cast<const float*>(s.h_out + kOutHead);
        const int32_t* h_lab = reinterpret_cast<const int32_t*>(s.h_out + s.off_lab);
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t b = 0; b < s.tickets.size(); ++b) {
            Request& r = requests_[s.tickets[b] - first_ticket_];
            const size_t row = b * static_cast<size_t>(cfg_.k_max);
            r.dist.assign(h_dist + row, h_dist + row + r.k);
            r.lab.assign(h_lab + row, h_lab + row + r.k);
            r.done_ns = t_done;
        }
        stats_.completed += static_cast<int64_t>(s.tickets.size());
        inflight_bytes_ -= s.bytes;
        s.busy = false;
        s.tickets.clear();
        any = true;
        cv_done_.notify_all();
    }
    return any;
}

// Oldest pending requests into free streams, within the scratch budget.
bool Server::admit() {
    bool any = false;
    for (Slot& s : slots_) {
        if (s.busy) continue;
        std::vector<int64_t> batch;
        int64_t busy = 1;  // batches in flight, this one included
        int64_t load = 0;  // queries in flight, this batch's included
        bool big = false;  // a batch of more than 64 queries is in flight
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (pending_.empty()) return any;
            // the oldest requests that fit the batch size and the scratch budget
            int64_t bytes = 0, nnz = 0;
            size_t n = 0;
            while (n < pending_.size() && static_cast<int32_t>(n) < cfg_.max_batch) {
                const Request& r = requests_[pending_[n] - first_ticket_];
                const int64_t need = request_bytes(static_cast<int32_t>(r.indices.size()));
                if (inflight_bytes_ + bytes + need > cfg_.scratch_budget) {
                    ++stats_.deferred_by_budget;
                    break;
                }
                nnz += static_cast<int64_t>(r.indices.size());
                bytes += need;
                ++n;
            }
            if (n == 0) return any;
            // pack them into this stream's pinned input block, zero padded
            s.in_bytes = input_layout(static_cast<int64_t>(n), nnz, &s.off_ks, &s.off_sel, &s.off_ix, &s.off_v);
            std::memset(s.h_in, 0, s.off_ks);
            std::memset(s.h_in + s.off_ks, 0, s.off_ix - s.off_ks);
            std::memset(s.h_in + s.off_ix + nnz * 2, 0, s.off_v - s.off_ix - nnz * 2);
            std::memset(s.h_in + s.off_v + nnz * 4, 0, kPad);
            int32_t* h_ip = reinterpret_cast<int32_t*>(s.h_in);
            int32_t* h_ks = reinterpret_cast<int32_t*>(s.h_in + s.off_ks);
            int32_t* h_sel = reinterpret_cast<int32_t*>(s.h_in + s.off_sel);
            uint16_t* h_ix = reinterpret_cast<uint16_t*>(s.h_in + s.off_ix);
            float* h_v = reinterpret_cast<float*>(s.h_in + s.off_v);
            int64_t at = 0;
            for (size_t b = 0; b < n; ++b) {
                const Request& r = requests_[pending_.front() - first_ticket_];
                h_ip[b] = static_cast<int32_t>(at);
                std::memcpy(h_ix + at, r.indices.data(), r.indices.size() * 2);
                std::memcpy(h_v + at, r.values.data(), r.values.size() * 4);
                h_ks[b] = r.k;
                h_sel[b] = r.bitmap_id;
                at += static_cast<int64_t>(r.indices.size());
                batch.push_back(pending_.front());
                pending_.pop_front();
            }
            h_ip[n] = static_cast<int32_t>(at);
            s.bytes = bytes;
            inflight_bytes_ += bytes;
            ++stats_.batches;
            stats_.max_inflight_bytes = std::max(stats_.max_inflight_bytes, inflight_bytes_);
            stats_.max_batch_seen = std::max<int64_t>(stats_.max_batch_seen, static_cast<int64_t>(batch.size()));
            load = static_cast<int64_t>(batch.size());
            for (const Slot& o : slots_) {
                busy += o.busy ? 1 : 0;
                load += o.busy ? static_cast<int64_t>(o.tickets.size()) : 0;
                big = big || (o.busy && o.tickets.size() > 64);
            }
            stats_.max_concurrent_batches = std::max(stats_.max_concurrent_batches, busy);
        }
        launch(s, batch, big || load > spread_load_ ? 8 : 64);
        any = true;
    }
    return any;
}

// The spread search puts each query of a batch on the whole GPU, about 0.35 ms
// plus 13 us per query at 8.8M docs, against about 1.5 ms on one CU per query
// for the per-query workgroup. It costs more GPU time per query, so batches of
// up to 64 take it under light load, batches of up to 8 under heavy load:
// more than spread_load_ queries in flight (this batch's included), or a
// batch of more than 64 in flight.
void Server::launch(Slot& s, const std::vector<int64_t>& batch, int64_t spread_cap) {
    const int64_t n = static_cast<int64_t>(batch.size());
    hip_check(hipMemcpyAsync(s.d_in.get(), s.h_in, s.in_bytes, hipMemcpyHostToDevice, s.stream), "H2D");
    const size_t out = static_cast<size_t>(n) * cfg_.k_max;
    s.off_lab = kOutHead + up256(out * 4);
    s.out_bytes = s.off_lab + out * 4;
    char* d_in = s.d_in.get<char>();
    char* d_out = s.d_out.get<char>();
    const bool with_sel = !bitmaps_.empty();
    index_.search_device(n, reinterpret_cast<const int32_t*>(d_in), reinterpret_cast<const uint16_t*>(d_in + s.off_ix),
                         reinterpret_cast<const float*>(d_in + s.off_v),
                         reinterpret_cast<const int32_t*>(d_in + s.off_ks), cfg_.k_max, cfg_.req,
                         with_sel ? reinterpret_cast<const int32_t*>(d_in + s.off_sel) : nullptr,
                         with_sel ? d_bitmap_table_.get<const uint32_t*>() : nullptr,
                         static_cast<int32_t>(bitmaps_.size()), reinterpret_cast<float*>(d_out + kOutHead), reinterpret_cast<int32_t*>(d_out + s.off_lab),
                         nullptr, s.stream, reinterpret_cast<int32_t*>(d_out), spread_cap);
    // the error flag, the distances and the labels in one copy
    hip_check(hipMemcpyAsync(s.h_out, d_out, s.out_bytes, hipMemcpyDeviceToHost, s.stream), "D2H");
    hip_check(hipEventRecord(s.done, s.stream), "hipEventRecord");
    s.tickets = batch;
    s.busy = true;
}

}  // namespace nsparse_amd
