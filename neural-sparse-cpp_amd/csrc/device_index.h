// HBM layout of a SEISMIC / SEISMIC-SQ index and the argument blocks of the
// search and build kernels. Shared by host (.cpp) and device (.hip) code.
//
// Forward index: CSR exactly as the reference stores it (sparse_vectors.h:105-110):
//   indptr int32[N+1], indices uint16[nnz], values uint8[nnz] (SQ) or float[nnz].
// Posting lists, all `dim` lists concatenated (list l = term l):
//   list_cluster_base int32[L+1]  -> global cluster ids of list l
//   cluster_doc_ptr   int64[C+1]  -> docs of global cluster c (the reference's
//                                    per-list offsets_, rebased)
//   docs              int32[D]    -> doc ids, in the reference's docs_ order
//   list_term_base    int32[L+1]  -> summary columns of list l
//   sum_term_ids      uint16[T]   -> distinct summary terms of each list, ascending
//   sum_term_ptr      int64[T+1]  -> entries of column t (term_ptr_, rebased)
//   sum_cluster       uint16[E]   -> local cluster id per entry (csc_cluster_)
//   sum_value         uint8/float[E] (csc_value_)
// Block-contiguous row copies (uint8 indexes, optional): for posting position p
// (the order of `docs`), row_ref[p] packs the row's first word, its word count
// and its nnz (blk_ref_* below); the row of docs[p] is re-encoded into 16-byte
// words [first word, first word + word count) of blk_rows, so one cluster's
// rows are one contiguous, line-aligned stream. A word is 8 slots of 16 bits
// (slot s = bits 16s..16s+15), each slot (offset << 4) | code with a code of
// 1..15; an unused slot is 0:
//   * low-term word (slot 0's code nonzero): term = offset; terms < 4096;
//   * based word (slot 0's code zero): slot 0 holds the base term (a multiple
//     of 16), slots 1..7 hold term = base + offset, offsets <= 4095.
// An element whose code exceeds 15 becomes several slots of one term whose
// codes add up to it, and an element of code 0 none: the integer score
// sum(code * query code) is unchanged. A row's low terms come first, 8 to a
// word, then its other terms in row order, a new based word starting whenever
// a term falls outside [base, base + 4095] or 7 slots are full. A row holds an
// even number of words (a zero word completes an odd count), so every row
// starts on a 32-byte pair and is read as whole pairs.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace nsparse_amd {

// row_ref fields: nnz (bits 0-14), word count (bits 15-29), first word (bits 30-63).
constexpr int kBlkMaxField = 0x7FFF;
__host__ __device__ inline int64_t blk_ref(int64_t first_word, int words, int nnz) {
    return (first_word << 30) | (static_cast<int64_t>(words) << 15) | nnz;
}
__host__ __device__ inline int64_t blk_ref_first(int64_t r) { return r >> 30; }
__host__ __device__ inline int blk_ref_words(int64_t r) { return static_cast<int>((r >> 15) & kBlkMaxField); }
__host__ __device__ inline int blk_ref_nnz(int64_t r) { return static_cast<int>(r & kBlkMaxField); }
// Low-term words hold terms below this; based-word offsets stay below it.
constexpr int kBlkTermSpan = 4096;

struct DevIndexView {
    int32_t n_docs = 0;
    int32_t dim = 0;
    int32_t esz = 1;  // 1 = uint8 codes, 4 = float
    int32_t n_lists = 0;
    int32_t max_clusters = 0;  // largest cluster count of any list
    const int32_t* indptr = nullptr;
    const uint16_t* indices = nullptr;
    const void* values = nullptr;
    const int32_t* list_cluster_base = nullptr;
    const int64_t* cluster_doc_ptr = nullptr;
    const int32_t* docs = nullptr;
    const int32_t* list_term_base = nullptr;
    const uint16_t* sum_term_ids = nullptr;
    const int64_t* sum_term_ptr = nullptr;
    const uint16_t* sum_cluster = nullptr;
    const void* sum_value = nullptr;
    const int64_t* row_ref = nullptr;
    const uint4* blk_rows = nullptr;
    int64_t blk_words = 0;  // words in the block copies; word blk_words is zero padding
    // Per-list summary-term bitmaps (optional): bit t of list l's sum_nw words
    // is set when t is one of its summary terms; sum_rank[l * sum_nw + w] counts
    // the list's summary terms in words < w, so a term's column is
    // list_term_base[l] + sum_rank + the set bits below it in its word.
    const uint32_t* sum_bits = nullptr;
    const uint16_t* sum_rank = nullptr;
    int32_t sum_nw = 0;
};

// Per-query traversal counters (the algorithmic-byte model): one row of
// kStatFields int64 per query when SearchArgs::stats is set.
enum StatField {
    kStatLists = 0,         // cut lists visited (term < n_lists)
    kStatTermsFound,        // query entries found in a list's summary columns
    kStatCscEntries,        // summary (cluster, value) entries read
    kStatClustersTested,    // prune tests evaluated
    kStatClustersOpened,    // clusters whose doc ids were read
    kStatDocsRead,          // doc ids read from opened clusters
    kStatDocsScored,        // forward-index rows gathered and scored
    kStatNnzScored,         // non-zeros of those rows
    kStatProbeBytes,        // sum over visited lists of q_len * 2 * ceil(log2 T): the
                            // reference's binary-search probes over term_ids
    // wall-clock ticks (device wall clock) spent per phase, first lane's view
    kStatTicksSetup,        // query into LDS + cut
    kStatTicksSummary,      // summary columns -> block scores
    kStatTicksOrder,        // first-list cluster order
    kStatTicksPlan,         // chunk planning (wave 0)
    kStatTicksGather,       // doc ids, selector, row ranges
    kStatTicksScore,        // row gathers and dot products
    kStatTicksReplay,       // exact replay + top-k merge
    kStatFields
};

// What a device build did, for its byte count and rates.
struct BuildStats {
    int64_t kept_postings = 0;    // postings kept by lambda pruning
    int64_t kept_row_nnz = 0;     // non-zeros of those postings' doc rows
    int64_t clusters = 0;
    int64_t summary_entries = 0;
    int64_t assign_pairs = 0;     // (doc term, centroid posting) pairs scored by the assignment
};

struct SearchArgs {
    DevIndexView ix;
    const int32_t* q_indptr;   // device, n+1 (absolute offsets)
    const uint16_t* q_indices;
    const float* q_values;
    const int32_t* ks;         // device, per-query k (nullptr: k_max for all)
    int32_t k_max;             // row stride of the outputs
    int32_t cut;
    float heap_factor;
    float q_vmin;              // SQ query quantizer: code = lround(clamp((v - q_vmin) * q_scale))
    float q_scale;             // 255 / (vmax - vmin), computed on the host as the reference does
    float decode_scale;        // (ingest_range * query_range) / (255 * 255)
    const int32_t* sel_of_query;        // device, per-query selector slot or -1 (nullptr: use sel_all)
    const uint32_t* const* sel_bitmaps; // device array of device bitmap pointers
    int32_t n_bitmaps;                  // entries of sel_bitmaps; a slot >= n_bitmaps is rejected
    const uint32_t* sel_all;            // one bitmap for every query (nullptr: none)
    float* out_d;              // device, n x k_max
    int32_t* out_l;
    int64_t* stats;            // device, n x kStatFields, or nullptr
    int32_t* error_flag;       // device; set to a nonzero code on a rejected query
    int64_t n_queries;
    int64_t query_base;        // first query handled by this launch (one-workgroup-per-query kernel)
    int32_t spin_max = 1 << 22;  // polls before a wait between waves gives up (kDevChunkWait)
    int32_t* queue;            // persistent kernel: device counter, zero at launch
    const int32_t* order;      // persistent kernel: query of each counter value, or nullptr (identity)
};

// How the batched search runs (uint8 indexes; FP32 always uses the bitmap
// lookup on 256-lane workgroups). Every choice computes identical results.
struct SearchConfig {
    int32_t lookup = 1;         // 1: dense uint8 code table in LDS, 0: bitmap + rank table
    int32_t block = 512;        // lanes per query workgroup: 256 or 512
    int32_t persistent = 0;     // 1: persistent workgroups pulling queries from a device queue
    int32_t longest_first = 0;  // persistent only: hand out queries by descending non-zeros
};

// Error codes a kernel can leave in SearchArgs::error_flag.
enum DeviceError : int32_t {
    kDevOk = 0,
    kDevQueryTooLong = 1,   // query nnz above the device limit
    kDevTermOutOfRange = 2, // query term id >= dim
    kDevTooManyClusters = 3, // list with more clusters than the LDS plan allows
    kDevBadK = 4,           // per-query k outside [1, k_max]
    kDevChunkWait = 5,      // a wait between the waves of a search timed out
    kDevBadSelector = 6     // per-query selector slot >= the bitmap table's size
};

}  // namespace nsparse_amd
