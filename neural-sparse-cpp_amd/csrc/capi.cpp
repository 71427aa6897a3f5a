// extern "C" boundary: status codes + thread-local last error, no exception
// crosses it. Each function forwards to the C++ index family in index.cpp.
#include <sys/mman.h>
#include <sys/stat.h>
#include <fcntl.h>
#include <unistd.h>

#include <array>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_util.h"
#include "index.h"
#include "server.h"
#include "io.h"
#include "kernels.h"
#include "nsparse_c.h"

namespace nsparse_amd {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const char* msg) { g_last_error = msg; }
}  // namespace nsparse_amd

using namespace nsparse_amd;

struct nsparse_index {
    std::unique_ptr<Index> impl;
};

namespace {

Index& unwrap(const nsparse_index* h) {
    if (h == nullptr || !h->impl) throw std::invalid_argument("index handle is null");
    return *h->impl;
}

SeismicIndex& seismic_of(nsparse_index* h) {
    Index& ix = unwrap(h);
    if (auto* s = dynamic_cast<SeismicIndex*>(&ix)) return *s;
    if (auto* m = dynamic_cast<IDMapIndex*>(&ix))
        if (auto* s = dynamic_cast<SeismicIndex*>(m->delegate())) return *s;
    throw std::invalid_argument("not a SEISMIC index");
}

// Reads or maps a whole file (the reference's FileIOReader / MmapFile).
struct FileBytes {
    std::vector<uint8_t> buf;
    void* map = nullptr;
    size_t size = 0;
    FileBytes(const char* path, bool use_mmap) {
        if (path == nullptr) throw std::invalid_argument("unexpected nullptr");
        if (use_mmap) {
            const int fd = ::open(path, O_RDONLY);
            if (fd < 0) throw std::runtime_error(std::string("Cannot open file: ") + path);
            struct stat st {};
            if (::fstat(fd, &st) != 0) {
                ::close(fd);
                throw std::runtime_error(std::string("Cannot stat file: ") + path);
            }
            size = static_cast<size_t>(st.st_size);
            if (size > 0) {
                map = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
                if (map == MAP_FAILED) {
                    map = nullptr;
                    ::close(fd);
                    throw std::runtime_error(std::string("Cannot map file: ") + path);
                }
            }
            ::close(fd);
        } else {
            std::ifstream f(path, std::ios::binary | std::ios::ate);
            if (!f) throw std::runtime_error(std::string("Cannot open file: ") + path);
            size = static_cast<size_t>(f.tellg());
            f.seekg(0);
            buf.resize(size);
            if (size > 0 && !f.read(reinterpret_cast<char*>(buf.data()), static_cast<std::streamsize>(size)))
                throw std::runtime_error(std::string("Cannot read file: ") + path);
        }
    }
    ~FileBytes() {
        if (map != nullptr) ::munmap(map, size);
    }
    const uint8_t* data() const { return map != nullptr ? static_cast<const uint8_t*>(map) : buf.data(); }
};

template <class... P>
void throw_if_null_all(P*... ptrs) {
    if (((ptrs == nullptr) || ...)) throw std::invalid_argument("unexpected nullptr");
}

}  // namespace

struct nsparse_server {
    std::unique_ptr<Server> impl;
};

namespace {

// cut / heap_factor / SQ range of a device search (the selector field is not used)
SearchRequest device_request(const nsparse_search_params* params) {
    SearchRequest req;
    if (params != nullptr) {
        if (params->kind == NSPARSE_PARAMS_SEISMIC || params->kind == NSPARSE_PARAMS_SEISMIC_SQ) {
            req.cut = params->cut;
            req.heap_factor = params->heap_factor;
        }
        if (params->kind == NSPARSE_PARAMS_SEISMIC_SQ) {
            req.has_sq_range = true;
            req.vmin = params->vmin;
            req.vmax = params->vmax;
        }
    }
    return req;
}

SeismicIndex& device_search_index(nsparse_index* index) {
    if (dynamic_cast<IDMapIndex*>(&unwrap(index)) != nullptr)
        throw std::invalid_argument("device searches take the SEISMIC index itself, not an IDMap");
    return seismic_of(index);
}

}  // namespace

namespace {

Server& server_of(nsparse_server* s) {
    if (s == nullptr || !s->impl) throw std::invalid_argument("server handle is null");
    return *s->impl;
}

}  // namespace

extern "C" {

const char* nsparse_last_error(void) { return g_last_error.c_str(); }

int nsparse_index_factory(int dim, const char* description, nsparse_index** out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        auto h = std::make_unique<nsparse_index>();
        h->impl = index_factory(dim, description);
        *out = h.release();
    });
}

int nsparse_free(nsparse_index* index) {
    return guarded([&] { delete index; });
}

int nsparse_idmap_wrap(nsparse_index* delegate, nsparse_index** out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        unwrap(delegate);  // a live index
        auto h = std::make_unique<nsparse_index>();
        h->impl = std::make_unique<IDMapIndex>(std::move(delegate->impl));
        *out = h.release();
    });
}

int nsparse_index_id(const nsparse_index* index, char out_fourcc[4]) {
    return guarded([&] {
        if (out_fourcc == nullptr) throw std::invalid_argument("unexpected nullptr");
        const auto id = unwrap(index).id();
        std::memcpy(out_fourcc, id.data(), 4);
    });
}

int nsparse_get_dimension(const nsparse_index* index, int* out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out = unwrap(index).dimension();
    });
}

int nsparse_num_vectors(const nsparse_index* index, int64_t* out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out = unwrap(index).num_vectors();
    });
}

int nsparse_add(nsparse_index* index, nsparse_idx_t n, const nsparse_idx_t* indptr, const nsparse_term_t* indices,
                const float* values) {
    return guarded([&] { unwrap(index).add(n, indptr, indices, values); });
}

int nsparse_add_with_ids(nsparse_index* index, nsparse_idx_t n, const nsparse_idx_t* indptr,
                         const nsparse_term_t* indices, const float* values, const nsparse_idx_t* ids) {
    return guarded([&] { unwrap(index).add_with_ids(n, indptr, indices, values, ids); });
}

int nsparse_build(nsparse_index* index) {
    return guarded([&] { unwrap(index).build(); });
}

int nsparse_search(nsparse_index* index, nsparse_idx_t n, const nsparse_idx_t* indptr, const nsparse_term_t* indices,
                   const float* values, int k, float* distances, nsparse_idx_t* labels,
                   const nsparse_search_params* params) {
    return guarded([&] { unwrap(index).search(n, indptr, indices, values, k, distances, labels, params); });
}

int nsparse_search_device(nsparse_index* index, int64_t n, const int32_t* d_indptr, const uint16_t* d_indices,
                          const float* d_values, const int32_t* d_ks, int k_max, const nsparse_search_params* params,
                          const int32_t* d_selector_of_query, const uint32_t* const* d_bitmaps, int n_bitmaps,
                          float* d_distances, int32_t* d_labels, void* stream) {
    return guarded([&] {
        SeismicIndex& s = device_search_index(index);
        if (d_selector_of_query != nullptr && (d_bitmaps == nullptr || n_bitmaps <= 0))
            throw std::invalid_argument("per-query selectors need a bitmap table");
        s.search_device(n, d_indptr, d_indices, d_values, d_ks, k_max, device_request(params), d_selector_of_query,
                        d_bitmaps, n_bitmaps, d_distances, d_labels, nullptr, static_cast<hipStream_t>(stream));
    });
}

int nsparse_search_device_noalloc(nsparse_index* index, int64_t n, const int32_t* d_indptr,
                                  const uint16_t* d_indices, const float* d_values, const int32_t* d_ks, int k_max,
                                  const nsparse_search_params* params, const int32_t* d_selector_of_query,
                                  const uint32_t* const* d_bitmaps, int n_bitmaps, float* d_distances,
                                  int32_t* d_labels, int32_t* d_error, void* stream) {
    return guarded([&] {
        SeismicIndex& s = device_search_index(index);
        s.search_device_noalloc(n, d_indptr, d_indices, d_values, d_ks, k_max, device_request(params),
                                d_selector_of_query, d_bitmaps, n_bitmaps, d_distances, d_labels, d_error,
                                static_cast<hipStream_t>(stream));
    });
}

struct nsparse_search_plan {
    std::unique_ptr<SearchPlan> impl;
};

int nsparse_search_plan_create(nsparse_index* index, int64_t n, const int32_t* d_indptr, const uint16_t* d_indices,
                               const float* d_values, const int32_t* d_ks, int k_max,
                               const nsparse_search_params* params, const int32_t* d_selector_of_query,
                               const uint32_t* const* d_bitmaps, int n_bitmaps, float* d_distances,
                               int32_t* d_labels, const nsparse_search_config* config, void* stream,
                               nsparse_search_plan** out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out = nullptr;
        SeismicIndex& s = device_search_index(index);
        if (d_selector_of_query != nullptr && (d_bitmaps == nullptr || n_bitmaps <= 0))
            throw std::invalid_argument("per-query selectors need a bitmap table");
        require_device();
        const SearchArgs a = s.search_args(n, d_indptr, d_indices, d_values, d_ks, k_max, device_request(params),
                                           d_selector_of_query, d_bitmaps, n_bitmaps, d_distances, d_labels,
                                           nullptr);
        SearchConfig c;
        if (config != nullptr) {
            c.lookup = config->lookup;
            c.block = config->block;
            c.persistent = config->persistent;
            c.longest_first = config->longest_first;
        }
        auto plan = std::make_unique<nsparse_search_plan>();
        plan->impl = std::make_unique<SearchPlan>(s, a, k_max, config != nullptr ? &c : nullptr,
                                                  static_cast<hipStream_t>(stream));
        *out = plan.release();
    });
}

int nsparse_search_plan_launch(nsparse_search_plan* plan, void* stream) {
    return guarded([&] {
        if (plan == nullptr || !plan->impl) throw std::invalid_argument("search plan handle is null");
        plan->impl->launch(static_cast<hipStream_t>(stream));
    });
}

int nsparse_search_plan_config(const nsparse_search_plan* plan, nsparse_search_config* out, int max_candidates,
                               nsparse_search_config* candidates, float* ms, int* n_candidates) {
    return guarded([&] {
        if (plan == nullptr || !plan->impl) throw std::invalid_argument("search plan handle is null");
        const auto to_c = [](const SearchConfig& c) {
            return nsparse_search_config{c.lookup, c.block, c.persistent, c.longest_first};
        };
        if (out != nullptr) *out = to_c(plan->impl->config());
        const auto& t = plan->impl->tuning();
        if (n_candidates != nullptr) *n_candidates = static_cast<int>(t.size());
        for (int i = 0; i < max_candidates && i < static_cast<int>(t.size()); ++i) {
            if (candidates != nullptr) candidates[i] = to_c(t[i].first);
            if (ms != nullptr) ms[i] = t[i].second;
        }
    });
}

int nsparse_search_plan_free(nsparse_search_plan* plan) {
    return guarded([&] { delete plan; });
}

// Extension used by tests and the benchmark: same as nsparse_search_device plus
// per-query traversal counters (n x 8 int64, device) and the kernel error code.
int nsparse_search_device_stats(nsparse_index* index, int64_t n, const int32_t* d_indptr, const uint16_t* d_indices,
                                const float* d_values, const int32_t* d_ks, int k_max,
                                const nsparse_search_params* params, float* d_distances, int32_t* d_labels,
                                int64_t* d_stats, void* stream) {
    return guarded([&] {
        SeismicIndex& s = seismic_of(index);
        s.search_device(n, d_indptr, d_indices, d_values, d_ks, k_max, device_request(params), nullptr, nullptr, 0,
                        d_distances, d_labels, d_stats, static_cast<hipStream_t>(stream));
    });
}

int nsparse_server_create(nsparse_index* index, const nsparse_server_config* config, nsparse_server** out) {
    return guarded([&] {
        if (config == nullptr || out == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out = nullptr;
        SeismicIndex& s = device_search_index(index);
        ServerConfig c;
        c.max_batch = config->max_batch;
        c.n_streams = config->n_streams;
        c.scratch_budget = config->scratch_budget;
        c.k_max = config->k_max;
        c.req.cut = config->cut;
        c.req.heap_factor = config->heap_factor;
        auto h = std::make_unique<nsparse_server>();
        h->impl = std::make_unique<Server>(s, c);
        *out = h.release();
    });
}

int nsparse_server_add_bitmap(nsparse_server* server, const uint32_t* words, int64_t n_words, int32_t* out_id) {
    return guarded([&] {
        if (out_id == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out_id = server_of(server).add_bitmap(words, n_words);
    });
}

int nsparse_server_submit(nsparse_server* server, int32_t nnz, const uint16_t* indices, const float* values,
                          int32_t k, int32_t bitmap_id, int64_t arrival_ns, int64_t* out_ticket) {
    return guarded([&] {
        if (out_ticket == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out_ticket = server_of(server).submit(nnz, indices, values, k, bitmap_id, arrival_ns);
    });
}

int nsparse_server_wait(nsparse_server* server, int64_t ticket, float* distances, int32_t* labels,
                        double* latency_us) {
    return guarded([&] {
        const double us = server_of(server).wait(ticket, distances, labels);
        if (latency_us != nullptr) *latency_us = us;
    });
}

int nsparse_server_now_ns(nsparse_server* server, int64_t* out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out = server_of(server).now_ns();
    });
}

int nsparse_server_replay(nsparse_server* server, int64_t n, const double* arrival_s, const int32_t* indptr,
                          const uint16_t* indices, const float* values, const int32_t* ks,
                          const int32_t* bitmap_ids, float* distances, int32_t* labels, double* latency_us) {
    return guarded([&] {
        Server& srv = server_of(server);
        if (n <= 0) throw std::invalid_argument("value must be positive");
        throw_if_null_all(arrival_s, indptr, ks, distances, labels, latency_us);
        const int32_t k_max = srv.k_max();
        std::vector<int64_t> tickets(static_cast<size_t>(n));
        const int64_t t0 = srv.now_ns();
        for (int64_t i = 0; i < n; ++i) {
            const int64_t at = t0 + static_cast<int64_t>(arrival_s[i] * 1e9);
            while (srv.now_ns() < at) {
            }
            const int32_t a = indptr[i], b = indptr[i + 1];
            tickets[i] = srv.submit(b - a, indices + a, values + a, ks[i], bitmap_ids ? bitmap_ids[i] : -1, at);
        }
        for (int64_t i = 0; i < n; ++i) {
            float* d = distances + i * k_max;
            int32_t* l = labels + i * k_max;
            std::fill(d, d + k_max, -1.0F);
            std::fill(l, l + k_max, -1);
            latency_us[i] = srv.wait(tickets[i], d, l);
        }
    });
}

int nsparse_server_stats(nsparse_server* server, int64_t* out6) {
    return guarded([&] {
        if (out6 == nullptr) throw std::invalid_argument("unexpected nullptr");
        const ServerStats st = server_of(server).stats();
        out6[0] = st.completed;
        out6[1] = st.batches;
        out6[2] = st.max_inflight_bytes;
        out6[3] = st.max_batch_seen;
        out6[4] = st.deferred_by_budget;
        out6[5] = st.max_concurrent_batches;
    });
}

int nsparse_server_free(nsparse_server* server) {
    return guarded([&] { delete server; });
}

int nsparse_device_error(nsparse_index* index, int32_t* out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        *out = seismic_of(index).last_device_error();
    });
}

int nsparse_build_stats(nsparse_index* index, int64_t* out5) {
    return guarded([&] {
        if (out5 == nullptr) throw std::invalid_argument("unexpected nullptr");
        const BuildStats& s = seismic_of(index).build_stats();
        out5[0] = s.kept_postings;
        out5[1] = s.kept_row_nnz;
        out5[2] = s.clusters;
        out5[3] = s.summary_entries;
        out5[4] = s.assign_pairs;
    });
}

int nsparse_exact_scores_device(int64_t n_docs, const int32_t* d_indptr, const uint16_t* d_indices,
                                const float* d_values, int dim, int q_len, const uint16_t* d_q_indices,
                                const float* d_q_values, float* d_scores, void* stream) {
    return guarded([&] {
        if (n_docs <= 0 || dim <= 0 || q_len < 0) throw std::invalid_argument("value must be positive");
        throw_if_null_all(d_indptr, d_indices, d_values, d_scores);
        if (q_len > 0) throw_if_null_all(d_q_indices, d_q_values);
        require_device();
        launch_exact_scores(d_indptr, d_indices, d_values, n_docs, dim, d_q_indices, d_q_values, q_len, d_scores,
                            static_cast<hipStream_t>(stream));
    });
}

int nsparse_rescore_device(int backward, int64_t n_queries, const int32_t* d_q_indptr, const uint16_t* d_q_indices,
                           const float* d_q_values, int n_cand, const int32_t* d_candidates, int64_t n_docs,
                           const int32_t* d_indptr, const uint16_t* d_indices, const float* d_values, int dim,
                           float* d_scores, const float* d_grad_scores, float* d_grad_q, float* d_grad_d,
                           int32_t* d_error, void* stream) {
    return guarded([&] {
        if (n_queries <= 0 || n_cand <= 0 || n_docs <= 0 || dim <= 0)
            throw std::invalid_argument("value must be positive");
        throw_if_null_all(d_q_indptr, d_q_indices, d_q_values, d_candidates, d_indptr, d_indices, d_values, d_error);
        if (!backward && d_scores == nullptr) throw std::invalid_argument("unexpected nullptr");
        if (backward && d_grad_scores == nullptr) throw std::invalid_argument("unexpected nullptr");
        require_device();
        launch_rescore(backward != 0, n_queries, d_q_indptr, d_q_indices, d_q_values, n_cand, d_candidates, n_docs,
                       d_indptr, d_indices, d_values, dim, d_scores, d_grad_scores, d_grad_q, d_grad_d, d_error,
                       static_cast<hipStream_t>(stream));
    });
}

int nsparse_write_index(nsparse_index* index, const char* path) {
    return guarded([&] {
        if (path == nullptr) throw std::invalid_argument("unexpected nullptr");
        std::FILE* f = std::fopen(path, "wb");
        if (f == nullptr) throw std::runtime_error(std::string("Cannot open file for writing: ") + path);
        try {
            ByteWriter w(f);
            write_index(unwrap(index), w);
        } catch (...) {
            std::fclose(f);
            throw;
        }
        if (std::fclose(f) != 0) throw std::runtime_error("closing the index file failed");
    });
}

int nsparse_write_index_buffer(nsparse_index* index, uint8_t** out_buf, size_t* out_size) {
    return guarded([&] {
        if (out_buf == nullptr || out_size == nullptr) throw std::invalid_argument("unexpected nullptr");
        ByteWriter w;
        write_index(unwrap(index), w);
        auto& b = w.buffer();
        *out_buf = static_cast<uint8_t*>(std::malloc(b.size() == 0 ? 1 : b.size()));
        if (*out_buf == nullptr) throw std::bad_alloc();
        std::memcpy(*out_buf, b.data(), b.size());
        *out_size = b.size();
    });
}

void nsparse_buffer_free(uint8_t* buf) { std::free(buf); }

int nsparse_read_index(const char* path, int flags, nsparse_index** out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("unexpected nullptr");
        FileBytes fb(path, (flags & NSPARSE_IO_FLAG_MMAP) != 0);
        ByteReader r(fb.data(), fb.size);
        auto h = std::make_unique<nsparse_index>();
        h->impl = read_index(r);
        *out = h.release();
    });
}

int nsparse_read_index_buffer(const uint8_t* buf, size_t size, int flags, nsparse_index** out) {
    return guarded([&] {
        (void)flags;
        if (out == nullptr || (buf == nullptr && size > 0)) throw std::invalid_argument("unexpected nullptr");
        ByteReader r(buf, size);
        auto h = std::make_unique<nsparse_index>();
        h->impl = read_index(r);
        *out = h.release();
    });
}

// Index::read_csr (index.cpp:63-142): int64 header (rows, cols, nnz), int64
// indptr, int32 indices, float32 values.
int nsparse_read_csr(nsparse_index* index, const char* path, int residency) {
    return guarded([&] {
        Index& ix = unwrap(index);
        if (path == nullptr) throw std::invalid_argument("file_path must not be null");
        if (residency == 1 && dynamic_cast<SeismicIndex*>(&ix) != nullptr &&
            static_cast<SeismicIndex&>(ix).sq())
            throw std::invalid_argument(
                "mmap residency is not available for a quantized index: a mapped CSR is borrowed as float, and "
                "this index searches over codes");
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f) throw std::invalid_argument(std::string("CSR file does not exist: ") + path);
        const auto fsize = static_cast<uint64_t>(f.tellg());
        f.seekg(0);
        int64_t hdr[3];
        if (!f.read(reinterpret_cast<char*>(hdr), sizeof(hdr)))
            throw std::runtime_error(std::string("Truncated CSR file: ") + path);
        const int64_t rows = hdr[0], cols = hdr[1], nnz = hdr[2];
        if (rows <= 0 || cols <= 0 || nnz < 0) throw std::invalid_argument(std::string("Invalid CSR header in: ") + path);
        if (rows > INT32_MAX || nnz > INT32_MAX)
            throw std::invalid_argument(std::string("CSR file too large for 32-bit offsets: ") + path);
        if (cols > ix.dimension())
            throw std::invalid_argument(std::string("CSR column count exceeds index dimension: ") + path);
        const uint64_t expect = 24 + static_cast<uint64_t>(rows + 1) * 8 + static_cast<uint64_t>(nnz) * 8;
        if (fsize != expect)
            throw std::invalid_argument(std::string("CSR file is not in the interchange layout: ") + path);
        std::vector<int64_t> ip64(rows + 1);
        f.read(reinterpret_cast<char*>(ip64.data()), static_cast<std::streamsize>(ip64.size() * 8));
        if (ip64.front() != 0 || ip64.back() != nnz)
            throw std::invalid_argument(std::string("Inconsistent CSR indptr in: ") + path);
        std::vector<int32_t> ix32(nnz);
        f.read(reinterpret_cast<char*>(ix32.data()), static_cast<std::streamsize>(nnz * 4));
        std::vector<float> vals(nnz);
        f.read(reinterpret_cast<char*>(vals.data()), static_cast<std::streamsize>(nnz * 4));
        if (!f) throw std::runtime_error(std::string("Truncated CSR file: ") + path);
        std::vector<int32_t> ip(ip64.begin(), ip64.end());
        std::vector<uint16_t> terms(nnz);
        for (int64_t i = 0; i < nnz; ++i) {
            if (ix32[i] < 0 || ix32[i] >= cols || ix32[i] > 65535)
                throw std::invalid_argument(std::string("CSR term out of range in: ") + path);
            terms[i] = static_cast<uint16_t>(ix32[i]);
        }
        ix.add(static_cast<int32_t>(rows), ip.data(), terms.data(), vals.data());
    });
}

int nsparse_get_vectors(const nsparse_index* index, int64_t* n_rows, int64_t* nnz, int* element_size,
                        nsparse_idx_t* indptr, nsparse_term_t* indices, uint8_t* values) {
    return guarded([&] {
        const HostCsr* v = unwrap(index).vectors();
        if (n_rows) *n_rows = v ? v->n() : 0;
        if (nnz) *nnz = v ? static_cast<int64_t>(v->indices.size()) : 0;
        if (element_size) *element_size = v ? static_cast<int>(v->esz) : 0;
        if (v == nullptr) return;
        if (indptr) std::memcpy(indptr, v->indptr.data(), v->indptr.size() * 4);
        if (indices) std::memcpy(indices, v->indices.data(), v->indices.size() * 2);
        if (values) std::memcpy(values, v->values.data(), v->values.size());
    });
}

}  // extern "C"

// ---- test hooks (include/nsparse_testing.h) ----
#include "nsparse_testing.h"
#include "stl_order.h"

extern "C" int nsparse_testing_cluster_order(const float* keys, int n, int64_t* out_order) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (keys == nullptr || out_order == nullptr)))
            throw std::invalid_argument("unexpected nullptr");
        std::vector<KeyItem> a(n);
        for (int i = 0; i < n; ++i) a[i] = KeyItem{keys[i], static_cast<uint32_t>(i)};
        introsort_desc(a.data(), n);
        for (int i = 0; i < n; ++i) out_order[i] = a[i].id;
    });
}

extern "C" int nsparse_testing_top_k_tokens(const uint16_t* terms, const float* weights, int n, int k, uint16_t* out,
                                            int* out_n) {
    return guarded([&] {
        if (k <= 0) throw std::invalid_argument("value must be positive");
        std::vector<KeyItem> heap(std::max(k, 1));
        *out_n = topk_tokens(
            n, k, [&](int i) { return weights[i]; }, [&](int i) { return static_cast<uint32_t>(terms[i]); },
            heap.data(), out);
    });
}

extern "C" int nsparse_testing_device_cluster_order(const float* keys, int n, int workgroup, int64_t* out_order) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (keys == nullptr || out_order == nullptr)))
            throw std::invalid_argument("unexpected nullptr");
        require_device();
        device_cluster_order(keys, n, out_order, workgroup != 0);
    });
}

extern "C" int nsparse_testing_read_probe(const void* d_buf, size_t bytes, uint32_t* d_out, void* stream) {
    return guarded([&] {
        require_device();
        device_read_probe(d_buf, bytes, d_out, static_cast<hipStream_t>(stream));
    });
}

extern "C" int nsparse_testing_device_top_k_tokens(const uint16_t* terms, const float* weights, int n, int k,
                                                   uint16_t* out, int* out_n) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (terms == nullptr || weights == nullptr || out == nullptr)) || out_n == nullptr)
            throw std::invalid_argument("unexpected nullptr");
        require_device();
        *out_n = device_top_k_tokens(terms, weights, n, k, out);
    });
}
