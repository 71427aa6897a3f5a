// Reusable staging for searches over host arrays.
//
// A search over host arrays needs a few device arrays (the queries, the
// outputs, a kernel's scratch) for the length of one call. Allocating and
// freeing them per call costs an allocator round trip each, and hipFree
// waits for the whole device, so concurrent callers would queue behind each
// other's kernels. A HostStage instead keeps one stream, one device arena and
// one pinned host mirror per concurrent caller: a call lays its inputs out in
// the mirror, moves them over in one copy, launches on the stage's stream,
// brings the outputs back in one copy and waits for that stream alone.
//
// The first stage of a pool runs on the null stream, which the process has
// anyway; only callers that overlap it get streams of their own. HIP maps a
// process's streams onto a few hardware queues (four by default), and one
// stream more, even idle, makes other streams share queues: a serving loop's
// slots then queue behind each other (p50 at 10k requests/s on the 8.8M-doc
// index went from 2.4 to 3.9 ms with one extra stream in the process).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <string>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "device_mem.h"

namespace nsparse_amd {

class HostStage {
public:
    enum Kind { kIn = 0, kOut = 1, kScratch = 2 };

    explicit HostStage(bool own_stream) {
        if (own_stream) hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
    HostStage(const HostStage&) = delete;
    HostStage& operator=(const HostStage&) = delete;
    ~HostStage() {
        if (dev_ != nullptr) (void)hipFree(dev_);
        if (host_ != nullptr) (void)hipHostFree(host_);
        if (stream_ != nullptr) (void)hipStreamDestroy(stream_);
    }

    hipStream_t stream() const { return stream_; }

    // Starts the layout of a call. Segments are laid out inputs first, then
    // outputs, then scratch, each 256-byte aligned and followed by at least 16
    // zero bytes on the input side (the padding that lets vector loads run
    // past an array's end).
    void begin() { segs_.clear(); }
    int add(Kind kind, size_t bytes) {
        segs_.push_back({kind, bytes, 0});
        return static_cast<int>(segs_.size()) - 1;
    }
    // Assigns offsets and grows the arena and the mirror when the call needs more.
    void commit() {
        size_t off = 0;
        for (int kind = kIn; kind <= kScratch; ++kind) {
            if (kind == kOut) in_end_ = off;
            if (kind == kScratch) out_end_ = off;
            for (Seg& s : segs_)
                if (s.kind == kind) {
                    s.off = off;
                    off = (off + s.bytes + kPad + 255) & ~size_t{255};
                }
        }
        const size_t dev_need = std::max<size_t>(off, 256);
        const size_t host_need = std::max<size_t>(out_end_, 256);
        if (dev_need > dev_bytes_) {
            if (dev_ != nullptr) hip_check(hipFree(dev_), "hipFree");
            dev_ = nullptr;
            dev_bytes_ = 0;
            hip_check(hipMalloc(&dev_, dev_need), "hipMalloc");
            dev_bytes_ = dev_need;
        }
        if (host_need > host_bytes_) {
            if (host_ != nullptr) hip_check(hipHostFree(host_), "hipHostFree");
            host_ = nullptr;
            host_bytes_ = 0;
            hip_check(hipHostMalloc(&host_, host_need, hipHostMallocDefault), "hipHostMalloc");
            host_bytes_ = host_need;
        }
        // the input padding travels with the inputs
        for (const Seg& s : segs_)
            if (s.kind == kIn) std::memset(static_cast<char*>(host_) + s.off + s.bytes, 0, kPad);
    }
    template <class T = void>
    T* host(int seg) const {
        return reinterpret_cast<T*>(static_cast<char*>(host_) + segs_[seg].off);
    }
    template <class T = void>
    T* dev(int seg) const {
        return reinterpret_cast<T*>(static_cast<char*>(dev_) + segs_[seg].off);
    }
    // Copies a host array into an input segment's mirror.
    void put(int seg, const void* src, size_t bytes) {
        if (bytes > 0) std::memcpy(host<char>(seg), src, bytes);
    }
    void upload() {
        if (in_end_ > 0)
            hip_check(hipMemcpyAsync(dev_, host_, in_end_, hipMemcpyHostToDevice, stream_), "hipMemcpyAsync H2D");
    }
    // Brings every output segment back and waits for the stage's stream.
    void download_and_wait() {
        if (out_end_ > in_end_)
            hip_check(hipMemcpyAsync(static_cast<char*>(host_) + in_end_, static_cast<char*>(dev_) + in_end_,
                                     out_end_ - in_end_, hipMemcpyDeviceToHost, stream_),
                      "hipMemcpyAsync D2H");
        hip_check(hipStreamSynchronize(stream_), "search");
    }
    // Gives back an arena or a mirror above the sizes worth keeping between calls.
    void trim() {
        if (dev_bytes_ > kKeepDevice) {
            (void)hipFree(dev_);
            dev_ = nullptr;
            dev_bytes_ = 0;
        }
        if (host_bytes_ > kKeepHost) {
            (void)hipHostFree(host_);
            host_ = nullptr;
            host_bytes_ = 0;
        }
    }

private:
    static constexpr size_t kPad = 16;
    static constexpr size_t kKeepDevice = size_t{256} << 20;
    static constexpr size_t kKeepHost = size_t{64} << 20;
    struct Seg {
        Kind kind;
        size_t bytes, off;
    };
    hipStream_t stream_ = nullptr;
    void* dev_ = nullptr;
    size_t dev_bytes_ = 0;
    void* host_ = nullptr;
    size_t host_bytes_ = 0;
    size_t in_end_ = 0, out_end_ = 0;
    std::vector<Seg> segs_;
};

// The stages of one index: a call takes a free one (or makes one) and gives
// it back when done, so concurrent callers never share a stage.
class HostStagePool {
public:
    class Lease {
    public:
        Lease(HostStagePool& pool, std::unique_ptr<HostStage> s) : pool_(pool), s_(std::move(s)) {}
        Lease(const Lease&) = delete;
        Lease& operator=(const Lease&) = delete;
        ~Lease() {
            // a call that threw may have left copies in flight on the stream
            (void)hipStreamSynchronize(s_->stream());
            s_->trim();
            pool_.give_back(std::move(s_));
        }
        HostStage* operator->() const { return s_.get(); }
        HostStage& operator*() const { return *s_; }

    private:
        HostStagePool& pool_;
        std::unique_ptr<HostStage> s_;
    };

    Lease take() {
        bool first = false;
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (!free_.empty()) {
                std::unique_ptr<HostStage> s = std::move(free_.back());
                free_.pop_back();
                return Lease(*this, std::move(s));
            }
            first = made_++ == 0 && null_stream_first();
        }
        return Lease(*this, std::make_unique<HostStage>(!first));
    }

private:
    // The null stream synchronizes with every blocking stream of the process
    // (a search waits for work queued there before it, and work queued there
    // after it waits for the search); NSPARSE_STAGE_NULL_STREAM=0 gives the
    // first stage a non-blocking stream of its own instead.
    static bool null_stream_first() {
        const char* e = std::getenv("NSPARSE_STAGE_NULL_STREAM");
        return !(e != nullptr && std::string(e) == "0");
    }
    // The null-stream stage goes to the back, where take() looks first.
    void give_back(std::unique_ptr<HostStage> s) {
        std::lock_guard<std::mutex> lock(mu_);
        if (s->stream() == nullptr)
            free_.push_back(std::move(s));
        else
            free_.insert(free_.begin(), std::move(s));
    }
    std::mutex mu_;
    std::vector<std::unique_ptr<HostStage>> free_;
    int made_ = 0;
};

}  // namespace nsparse_amd
