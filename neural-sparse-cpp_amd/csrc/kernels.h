// Host-callable entry points of the HIP translation units (search.hip, search_spread.hip,
// build.hip, rescore.hip, testing.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "device_index.h"
#include "device_mem.h"

namespace nsparse_amd {

// Longest query (non-zeros) a search workgroup holds in LDS.
constexpr int kDeviceQueryMax = 512;

// ExactMatcher (exact_matcher.h:23-58): score an explicit sorted id list.
struct ExactArgs {
    DevIndexView ix;
    const int32_t* q_indptr;
    const uint16_t* q_indices;
    const float* q_values;
    const int32_t* ids;  // sorted, may repeat (ArrayIDSelector)
    int32_t n_ids;       // <= k
    int32_t k;
    float q_vmin, q_scale, decode_scale;
    float* out_d;
    int32_t* out_l;
    int64_t n_queries;
};

// Posting lists produced by the GPU build, in the layout of device_index.h.
struct DeviceLists {
    DeviceBuffer list_cluster_base, cluster_doc_ptr, docs, list_term_base, sum_term_ids, sum_term_ptr, sum_cluster,
        sum_value, list_entry_base;
    int32_t n_lists = 0;
    int max_clusters = 0;
    BuildStats stats;
};

// ScalarQuantizer::encode of host values on the GPU (bits 8: one byte per code,
// 16: two, little-endian); codes come back to the host.
void encode_sq_on_device(const float* host_values, size_t n, float vmin, float vmax, uint8_t* host_codes,
                         int bits = 8);

// build_inverted_lists_clusters (seismic_common.h:146-184) on the GPU.
void build_on_device(const DevIndexView& fwd, int64_t n_docs, int64_t nnz, int lambda, int beta, float alpha,
                     int seed, DeviceLists& out);

// Block-contiguous row copies of every posting (uint8 indexes) for the device
// search: row_ref (D int64) and blk_rows (16-byte words, the layout in
// device_index.h). Returns false, leaving the buffers empty, when HBM cannot
// hold them, a row exceeds the row_ref fields or a list's rows exceed 2^28 words.
bool build_block_rows(const DevIndexView& ix, int64_t n_postings, DeviceBuffer& row_ref, DeviceBuffer& blk_rows,
                      int64_t& n_words);

// Per-list summary-term bitmaps and rank prefixes (device_index.h) of the
// lists in `ix`. Returns false, leaving the buffers empty, when HBM cannot hold
// them or a list has 65536 or more summary terms.
bool build_summary_bitmaps(const DevIndexView& ix, DeviceBuffer& bits, DeviceBuffer& rank);

// The default (NSPARSE_SEARCH_LOOKUP / NSPARSE_SEARCH_BLOCK override it).
SearchConfig default_search_config(bool sq);
// `c` made valid for this index: the dense table is dropped when it does not fit in LDS.
SearchConfig resolve_search_config(int dim, int max_clusters, int kmax, bool sq, SearchConfig c);
// Device scratch a persistent launch over n queries needs (queue counter, order).
size_t search_scratch_bytes(int64_t n);
// Launches the search of a.n_queries queries with `cfg`; only plain kernels are
// enqueued on `stream`, so the sequence can be captured in a hipGraph.
void launch_search(const SearchArgs& a, int max_clusters, int kmax, bool sq, const SearchConfig& cfg, void* scratch,
                   hipStream_t stream);
void launch_search_checked(const SearchArgs& a, int max_clusters, int kmax, bool sq, hipStream_t stream);
// The batched search with the persistent, longest-first schedule whose queue
// counter, histogram and query order live in a file-scope device array of the
// library (no allocation): at most kNoAllocMaxQueries queries per call, and
// calls that overlap on one GPU race on that array.
constexpr int64_t kNoAllocMaxQueries = int64_t{1} << 20;
void launch_search_noalloc(const SearchArgs& a, int max_clusters, int kmax, bool sq, hipStream_t stream);
void launch_exact_checked(const ExactArgs& a, bool sq, hipStream_t stream);
// The spread search (search_spread.hip): each query of a small batch on the
// whole GPU, one workgroup per (cut list, slice of its clusters), then an exact
// replay; uint8 indexes with block rows only, no traversal counters. Same
// results as launch_search. Scratch: search_spread_scratch_bytes, per launch.
// `qcap` is the longest query (non-zeros) the launch must take: its query
// tables are sized for max(qcap, kDeviceQueryMax), as far as LDS allows, so
// longer queries than the per-query workgroup's can run here.
bool search_spread_fits(const DevIndexView& ix, int max_clusters, int kmax, int cut, int64_t max_list_docs,
                        int qcap = kDeviceQueryMax);
size_t search_spread_scratch_bytes(const DevIndexView& ix, int max_clusters, int kmax, int cut,
                                   int64_t max_list_docs, int64_t n_queries, int qcap = kDeviceQueryMax);
void launch_search_spread(const SearchArgs& a, int max_clusters, int kmax, int64_t max_list_docs, void* scratch,
                          hipStream_t stream, int qcap = kDeviceQueryMax);
// Most docs any posting list of `ix` holds (device lists).
int64_t max_list_docs(const DevIndexView& ix);
// Test surface (testing.hip): the GPU's first-list order (one wavefront, or the
// whole-workgroup form) of at most kMaxTestOrder keys, and its query cut
// (topk_tokens_wave) of at most 512 entries with k <= 64; returns the count.
constexpr int kMaxTestOrder = 4096;
void device_cluster_order(const float* keys, int n, int64_t* out_order, bool block);
int device_top_k_tokens(const uint16_t* terms, const float* weights, int n, int k, uint16_t* out);
// One streaming read of `bytes` (a multiple of 16, 16-byte aligned) of device
// memory; *d_out receives the XOR of its 32-bit words. Measures the plain read rate.
void device_read_probe(const void* d_buf, size_t bytes, uint32_t* d_out, hipStream_t stream);
int search_lds_bytes(int dim, int max_clusters, int kmax, bool sq);
std::string device_error_message(int32_t code);
// Exhaustive search (BrutalIndex): every document scored, the k best kept per
// query (k <= kBruteKMax). Queries run in groups of brute_search_group() rows
// of n_docs float scores, which d_scores (brute_search_scratch_bytes) holds.
constexpr int kBruteKMax = 1024;
int64_t brute_search_group(int64_t n_docs, int64_t n_queries);
size_t brute_search_scratch_bytes(int64_t n_docs, int64_t n_queries);
void launch_brute_search(const int32_t* indptr, const uint16_t* indices, const float* values, int64_t n_docs, int dim,
                         const int32_t* d_q_indptr, const uint16_t* d_q_indices, const float* d_q_values,
                         int64_t n_queries, int k, float* d_out_d, int32_t* d_out_l, float* d_scores,
                         hipStream_t stream);
void launch_exact_scores(const int32_t* indptr, const uint16_t* indices, const float* values, int64_t n_docs, int dim,
                         const uint16_t* q_indices, const float* q_values, int qn, float* scores, hipStream_t stream);

// InvertedIndex posting lists on the device: list_ptr (int64, dim + 1), docs
// (int32) and vals (float) term-major with doc ids ascending inside a term,
// and each term's largest posting value from 0 up (float, dim).
struct DevicePostings {
    DeviceBuffer list_ptr, docs, vals, max_score;
    int64_t n_postings = 0;
    bool duplicate_docs = false;  // some row repeats a term, so a list holds that doc twice
    // Search table: window_start[t * (n_windows + 1) + w] is the position of
    // list t's first posting with doc id >= w * kInvertedWindow (int32; w =
    // n_windows gives the list's end). Made by build_window_starts.
    DeviceBuffer window_start;
    int64_t n_windows = 0;
};

// Doc ids per window of the exact inverted search (its LDS score tile).
constexpr int kInvertedWindow = 16384;

// The window_start table of `p` for lists over doc ids [0, n_docs).
void build_window_starts(DevicePostings& p, int64_t n_docs, int dim);

// ArrayInvertedLists::build_inverted_lists + InvertedIndex::build's per-term
// maxima (inverted_index.cpp:236-252) over a forward CSR already in HBM.
void build_postings_on_device(const int32_t* indptr, const uint16_t* indices, const float* values, int64_t n_docs,
                              int64_t nnz, int dim, DevicePostings& out);

// Exact InvertedIndex search: every posting of every query term is scored
// term by term into LDS windows of doc ids, then the k best positive scores
// (score desc, doc id asc) are kept, among the docs `allow` marks when it is
// not null (bit d of word d / 32). Scratch as for launch_brute_search.
void launch_inverted_search(const DevicePostings& p, int64_t n_docs, const int32_t* d_q_indptr,
                            const uint16_t* d_q_indices, const float* d_q_values, int64_t n_queries, int k,
                            const uint32_t* d_allow, float* d_out_d, int32_t* d_out_l, float* d_scores,
                            hipStream_t stream);

// Differentiable exact re-scoring (rescore.hip). Forward fills scores
// (n_q x n_cand); backward adds into grad_q / grad_d (either may be null).
void launch_rescore(bool backward, int64_t n_q, const int32_t* q_indptr, const uint16_t* q_indices,
                    const float* q_values, int n_cand, const int32_t* cand, int64_t n_docs, const int32_t* d_indptr,
                    const uint16_t* d_indices, const float* d_values, int dim, float* scores,
                    const float* grad_scores, float* grad_q, float* grad_d, int32_t* err, hipStream_t stream);

}  // namespace nsparse_amd
