/*
 * C ABI of the MI355X-native SEISMIC / SEISMIC-SQ index.
 *
 * Each entry point stands in for one member of the reference's nsparse::Index
 * surface (opensearch-project/neural-sparse-cpp, paths below are relative to
 * that repository's root). No C++ exception crosses this boundary: every call
 * returns an nsparse_status and leaves a message for nsparse_last_error().
 * The status codes map one-to-one onto the exception classes the reference
 * throws, so a C++ or SWIG adapter can re-throw the same types
 * (nsparse/utils/checks.h:19-84, nsparse/python/swignsparse.swig:47-59).
 *
 * Pointers are plain host pointers unless a name says "device". Output arrays
 * are caller-allocated, n*k, row-major, best first, padded with (-1.0f, -1)
 * exactly as the reference pads them (nsparse/seismic_index.cpp:270-273).
 */
#ifndef NSPARSE_C_H
#define NSPARSE_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t nsparse_idx_t;   /* nsparse/types.h:18 (idx_t)  */
typedef uint16_t nsparse_term_t; /* nsparse/types.h:19 (term_t) */
typedef struct nsparse_index nsparse_index;

typedef enum nsparse_status {
    NSPARSE_OK = 0,
    NSPARSE_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument -> ValueError   */
    NSPARSE_ERR_RUNTIME = 2,          /* std::runtime_error    -> RuntimeError */
    NSPARSE_ERR_OUT_OF_RANGE = 3,     /* std::out_of_range     -> IndexError   */
    NSPARSE_ERR_BAD_ALLOC = 4,        /* std::bad_alloc        -> MemoryError  */
    NSPARSE_ERR_DEVICE = 5,           /* HIP runtime failure (no reference twin) */
    NSPARSE_ERR_UNKNOWN = 6
} nsparse_status;

/* Message of the last failing call on this thread ("" if none). */
const char* nsparse_last_error(void);

/* index_factory(int dim, const char* desc) -- nsparse/index_factory.cpp:63-153.
 * Accepts the reference's strings unchanged: "seismic,lambda=..|beta=..|alpha=..|seed=..",
 * "seismic_sq,quantizer=8bit|vmin=..|vmax=..|lambda=..|beta=..|alpha=..|seed=..",
 * "brutal" (BrutalIndex, nsparse/brutal_index.cpp), "inverted" (InvertedIndex,
 * nsparse/inverted_index.cpp) and the "idmap,<delegate>" prefix. Unknown keys
 * are ignored, as there; "quantizer=16bit" selects 16-bit codes, any other
 * value 8-bit. */
int nsparse_index_factory(int dim, const char* description, nsparse_index** out);

/* delete index (SWIG %newobject ownership, nsparse/python/swignsparse.swig:182-184). */
int nsparse_free(nsparse_index* index);

/* IDMapIndex(Index*) -- nsparse/id_map_index.h:84, id_map_index.cpp:20: a new IDMap
 * index owning the delegate's index. `delegate` stays a handle (free it as usual)
 * that holds no index any more; calls on it report NSPARSE_ERR_INVALID_ARGUMENT. */
int nsparse_idmap_wrap(nsparse_index* delegate, nsparse_index** out);

/* Index::id() fourcc characters: SEIS / SESQ / IDMP / BRUT / INVT (nsparse/seismic_index.h:36,
 * nsparse/seismic_scalar_quantized_index.h:37, nsparse/id_map_index.h:81,
 * nsparse/brutal_index.h:32, nsparse/inverted_index.h:36). */
int nsparse_index_id(const nsparse_index* index, char out_fourcc[4]);

/* Index::get_dimension() -- nsparse/index.h:47. */
int nsparse_get_dimension(const nsparse_index* index, int* out);

/* Index::num_vectors() -- nsparse/index.h:51-54. */
int nsparse_num_vectors(const nsparse_index* index, int64_t* out);

/* Index::add -- nsparse/seismic_index.cpp:127-142, seismic_scalar_quantized_index.cpp:159-177,
 * brutal_index.cpp:27-42, inverted_index.cpp:218-234. Assumes indptr[0] == 0 like the
 * reference (sparse_vectors.cpp:92-126). Term ids must be < dim: checked here (the
 * reference does not check), except for "inverted", whose build() rejects them as the
 * reference's does ("term_id out of range"). */
int nsparse_add(nsparse_index* index, nsparse_idx_t n, const nsparse_idx_t* indptr,
                const nsparse_term_t* indices, const float* values);

/* Index::add_with_ids -- nsparse/id_map_index.cpp:63-75 (IDMapIndex only; others
 * report "add_with_ids not implemented in Index", index.cpp:57-60). */
int nsparse_add_with_ids(nsparse_index* index, nsparse_idx_t n, const nsparse_idx_t* indptr,
                         const nsparse_term_t* indices, const float* values, const nsparse_idx_t* ids);

/* Index::build -- nsparse/seismic_common.h:146-184 (SEISMIC) and
 * inverted_index.cpp:236-252 (inverted), run as HIP kernels; "brutal" has none
 * ("not implemented", as Index::build). */
int nsparse_build(nsparse_index* index);

/* ---- search parameters and selectors (nsparse/index.h:21-28, id_selector.h:22-86) ---- */

typedef enum nsparse_selector_kind {
    NSPARSE_SELECTOR_ARRAY = 0,  /* ArrayIDSelector(n, ids)  id_selector.h:51-66 */
    NSPARSE_SELECTOR_SET = 1,    /* SetIDSelector(n, ids)    id_selector.h:33-49 */
    NSPARSE_SELECTOR_NOT = 2,    /* NotIDSelector(inner)     id_selector.h:68-78 */
    NSPARSE_SELECTOR_BITMAP = 3,  /* membership bitmap over ids; not enumerable (extension) */
    NSPARSE_SELECTOR_CALLBACK = 4 /* is_member(ctx, id): any other IDSelector, id_selector.h:22-27;
                                     not enumerable (extension) */
} nsparse_selector_kind;

typedef struct nsparse_selector {
    int kind;                              /* nsparse_selector_kind */
    int64_t n;                             /* ARRAY/SET: id count; BITMAP: bit count */
    const nsparse_idx_t* ids;              /* ARRAY/SET */
    const uint32_t* bitmap;                /* BITMAP: bit i of word i/32 set => member */
    const struct nsparse_selector* inner;  /* NOT: non-owning, like the reference */
    /* CALLBACK: returns 1 for a member, 0 for a non-member, negative on failure
     * (the search then reports NSPARSE_ERR_RUNTIME). Called on the calling
     * thread, during the call only, with the ids the index holds (external ids
     * under an IDMap). */
    int (*is_member)(const void* ctx, nsparse_idx_t id);
    const void* ctx;
} nsparse_selector;

typedef enum nsparse_params_kind {
    NSPARSE_PARAMS_BASE = 0,      /* SearchParameters: selector only (index.h:21-28)      */
    NSPARSE_PARAMS_SEISMIC = 1,   /* SeismicSearchParameters(cut, heap_factor)           */
    NSPARSE_PARAMS_SEISMIC_SQ = 2 /* SeismicSQSearchParameters(vmin, vmax, cut, hf)      */
} nsparse_params_kind;

typedef struct nsparse_search_params {
    int kind;              /* nsparse_params_kind; BASE uses cut=10, heap_factor=1.0 */
    int cut;               /* seismic_index.h:27 default 10 */
    float heap_factor;     /* seismic_index.h:28 default 1.0 */
    float vmin, vmax;      /* SEISMIC_SQ: query quantizer range (seismic_scalar_quantized_index.cpp:182-191) */
    const nsparse_selector* selector; /* may be NULL */
} nsparse_search_params;

/* Index::search(n, indptr, indices, values, k, distances, labels, params) --
 * nsparse/index.cpp:32-49 (+ SeismicIndex::search seismic_index.cpp:152-220,
 * SeismicScalarQuantizedIndex::search seismic_scalar_quantized_index.cpp:201-305,
 * BrutalIndex::search brutal_index.cpp:44-96, InvertedIndex::search
 * inverted_index.cpp:254-293, IDMapIndex::search id_map_index.cpp:29-60).
 * params may be NULL. Safe to call from several threads on one built index. */
int nsparse_search(nsparse_index* index, nsparse_idx_t n, const nsparse_idx_t* indptr,
                   const nsparse_term_t* indices, const float* values, int k, float* distances,
                   nsparse_idx_t* labels, const nsparse_search_params* params);

/* Batched extension (no single reference twin: the reference API takes one k per
 * call). Queries and outputs are DEVICE pointers already resident in HBM; query i
 * asks for ks[i] results (ks may be NULL: all k_max), written to row i of an
 * n x k_max output and padded with (-1.0f, -1). selector_of_query[i] (may be NULL)
 * indexes bitmaps[] (n_bitmaps device pointers, NULL entry = no filter) or is -1;
 * an index >= n_bitmaps pads that query's row and reports device error 6
 * (nsparse_device_error). Runs as one launch (one workgroup per query) on `stream`
 * (hipStream_t, NULL = default) and does not synchronize. cut/heap_factor/vmin/vmax
 * come from params (selector field ignored). */
int nsparse_search_device(nsparse_index* index, int64_t n, const int32_t* d_indptr,
                          const uint16_t* d_indices, const float* d_values, const int32_t* d_ks,
                          int k_max, const nsparse_search_params* params,
                          const int32_t* d_selector_of_query, const uint32_t* const* d_bitmaps,
                          int n_bitmaps, float* d_distances, int32_t* d_labels, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). nsparse_search_device's search, computing identical results, with the
 * persistent longest-first schedule; its queue counter, histogram and query order
 * live in a static device array of the library (about 4 MiB per GPU, held until the
 * process ends), so the call allocates nothing. Limits: n <= 1048576 queries per
 * call (larger n: NSPARSE_ERR_INVALID_ARGUMENT); the index must already be resident
 * on the device (after nsparse_build, or one search of an index read from a file),
 * else NSPARSE_ERR_INVALID_ARGUMENT. The launch writes its device error code (0 =
 * none, else the codes nsparse_device_error reports) to d_error (device int32).
 * Calls that overlap on one GPU share the array and race: serialize them. */
int nsparse_search_device_noalloc(nsparse_index* index, int64_t n, const int32_t* d_indptr,
                                  const uint16_t* d_indices, const float* d_values, const int32_t* d_ks,
                                  int k_max, const nsparse_search_params* params,
                                  const int32_t* d_selector_of_query, const uint32_t* const* d_bitmaps,
                                  int n_bitmaps, float* d_distances, int32_t* d_labels, int32_t* d_error,
                                  void* stream);

/* nsparse_search_device plus per-query traversal counters: d_stats (device,
 * n x 16 int64) receives lists visited, query entries found in summaries,
 * summary entries read, prune tests, clusters opened, doc ids read, rows scored,
 * their non-zeros and the bytes of the reference's term_ids binary-search probes
 * (q_len * 2 * ceil(log2 T) per visited list) -- the terms of the
 * algorithmic-byte model -- then the device wall-clock ticks (100 MHz) spent in
 * setup, summary scoring, cluster ordering, chunk planning, gathering, scoring
 * and replay. Extension. */
int nsparse_search_device_stats(nsparse_index* index, int64_t n, const int32_t* d_indptr,
                                const uint16_t* d_indices, const float* d_values, const int32_t* d_ks,
                                int k_max, const nsparse_search_params* params, float* d_distances,
                                int32_t* d_labels, int64_t* d_stats, void* stream);

/* Schedule of a batched search (uint8 indexes; FP32 always runs lookup 0 on
 * 256-lane workgroups). Every choice computes identical results. Extension. */
typedef struct nsparse_search_config {
    int32_t lookup;        /* 1: dense uint8 code table in LDS, 0: bitmap + rank table */
    int32_t block;         /* lanes per query workgroup: 256, 512 or 1024 (1024: dense lookup only) */
    int32_t persistent;    /* 1: persistent workgroups pulling queries from a device queue */
    int32_t longest_first; /* persistent only: queries handed out by descending non-zeros */
} nsparse_search_config;

typedef struct nsparse_search_plan nsparse_search_plan;

/* The batched search of nsparse_search_device over fixed DEVICE buffers, captured
 * once in a hipGraph and replayed by nsparse_search_plan_launch. config NULL: the
 * autotuner times every valid schedule on this batch (on `stream`, synchronizing
 * it; the outputs are written meanwhile) and keeps the fastest. Between launches
 * the caller may rewrite the contents of the query and ks buffers (same sizes).
 * Free the plan before those buffers or the index. Extension. */
int nsparse_search_plan_create(nsparse_index* index, int64_t n, const int32_t* d_indptr,
                               const uint16_t* d_indices, const float* d_values, const int32_t* d_ks,
                               int k_max, const nsparse_search_params* params,
                               const int32_t* d_selector_of_query, const uint32_t* const* d_bitmaps,
                               int n_bitmaps, float* d_distances, int32_t* d_labels,
                               const nsparse_search_config* config, void* stream, nsparse_search_plan** out);
/* Replays the plan on `stream`; does not synchronize. */
int nsparse_search_plan_launch(nsparse_search_plan* plan, void* stream);
/* The schedule in use (*out), and the autotuner's timing of each candidate it
 * tried: up to max_candidates entries of candidates[] / ms[] (either may be NULL);
 * *n_candidates receives how many were timed (0 when the config was given). */
int nsparse_search_plan_config(const nsparse_search_plan* plan, nsparse_search_config* out,
                               int max_candidates, nsparse_search_config* candidates, float* ms,
                               int* n_candidates);
int nsparse_search_plan_free(nsparse_search_plan* plan);

/* ---- Serving: continuous batching over a device-resident index (extension) ----
 * Requests (one query each, own k, optional selector bitmap) are queued from any
 * thread; a dispatcher packs the oldest pending ones into batches, each run on one
 * of n_streams default-priority streams as one batched search. A batch is admitted
 * only while a stream is free and the HBM scratch of all batches in flight
 * (per request: 6 B per query non-zero + 12 B + 8 B * k_max) stays within
 * scratch_budget bytes; requests arriving meanwhile wait, then are admitted as
 * streams and budget free up, so batches overlap on the device. Each query's
 * results equal nsparse_search_device's for that query alone. */
typedef struct nsparse_server_config {
    int32_t max_batch;      /* queries per launch */
    int32_t n_streams;      /* batches in flight at most (1..16) */
    int64_t scratch_budget; /* bytes of in-flight request scratch admitted */
    int32_t k_max;          /* largest k a request may ask for */
    int32_t cut;
    float heap_factor;
} nsparse_server_config;

typedef struct nsparse_server nsparse_server;

int nsparse_server_create(nsparse_index* index, const nsparse_server_config* config, nsparse_server** out);
/* Host bitmap (bit d of word d/32 set = doc d allowed; at least ceil(n_docs/32)
 * words) copied into HBM; requests name it by *out_id. Register before submitting. */
int nsparse_server_add_bitmap(nsparse_server* server, const uint32_t* words, int64_t n_words, int32_t* out_id);
/* Queues one query (nnz <= 512, 1 <= k <= k_max, bitmap_id -1 = no filter).
 * arrival_ns < 0 stamps it now; else its arrival time in ns on the server clock
 * (nsparse_server_now_ns), from which its latency is measured. */
int nsparse_server_submit(nsparse_server* server, int32_t nnz, const uint16_t* indices, const float* values,
                          int32_t k, int32_t bitmap_id, int64_t arrival_ns, int64_t* out_ticket);
/* Blocks until the request is done; writes its k results (either pointer may be
 * NULL) and its latency (completion - arrival) in microseconds. */
int nsparse_server_wait(nsparse_server* server, int64_t ticket, float* distances, int32_t* labels,
                        double* latency_us);
int nsparse_server_now_ns(nsparse_server* server, int64_t* out);
/* Replays an arrival trace: query i (CSR rows of indptr/indices/values, ks[i],
 * bitmap_ids[i] or NULL) is submitted at arrival_s[i] seconds after the call (the
 * calling thread waits for each arrival time), then every result is collected into
 * row i of n x k_max distances/labels (padded (-1.0f, -1)) with its latency. */
int nsparse_server_replay(nsparse_server* server, int64_t n, const double* arrival_s, const int32_t* indptr,
                          const uint16_t* indices, const float* values, const int32_t* ks,
                          const int32_t* bitmap_ids, float* distances, int32_t* labels, double* latency_us);
/* completed, batches, max in-flight scratch bytes, largest batch, admission
 * rounds stopped by the budget, most batches in flight at once (6 int64). */
int nsparse_server_stats(nsparse_server* server, int64_t* out6);
int nsparse_server_free(nsparse_server* server);

/* Error code (0 = none) a search kernel left for the last nsparse_search_device*
 * call on this index: 1 query too long, 2 term id >= dim, 4 per-query k out of range.
 * Reading it synchronizes the device. Extension. */
int nsparse_device_error(nsparse_index* index, int32_t* out);

/* What the last nsparse_build of this SEISMIC index did (5 int64): postings kept by
 * lambda pruning, non-zeros of those postings' doc rows, clusters, summary entries,
 * and (doc term, centroid posting) pairs scored by the assignment. Zero for an index
 * read from a file. For the build's byte count and rates. Extension. */
int nsparse_build_stats(nsparse_index* index, int64_t* out5);

/* Exact scores of one query (host-free: all pointers DEVICE) against every row of
 * an FP32 CSR -- the scoring of BrutalIndex::single_query (nsparse/brutal_index.cpp:75-96),
 * used for recall ground truth. d_scores receives n_docs floats. Extension. */
int nsparse_exact_scores_device(int64_t n_docs, const int32_t* d_indptr, const uint16_t* d_indices,
                                const float* d_values, int dim, int q_len, const uint16_t* d_q_indices,
                                const float* d_q_values, float* d_scores, void* stream);

/* Differentiable exact re-scoring of candidates (training-side use of the scorer in
 * nsparse/seismic_common.h:78-103; no reference twin). All pointers DEVICE.
 * Forward (backward == 0): d_scores[q*n_cand + c] = sum_i qv[i] * doc(cand[q,c])[t_i].
 * Backward: d_grad_q[i] = sum_c g[q,c] * doc[t_i] (overwritten) and
 * d_grad_d[e] += g[q,c] * sum_{i: t_i = term(e)} qv[i] (accumulated; zero it first);
 * either gradient pointer may be NULL. cand < 0 marks an empty slot. Queries are
 * limited to 512 entries; a longer one sets *d_error to 1. Extension. */
int nsparse_rescore_device(int backward, int64_t n_queries, const int32_t* d_q_indptr,
                           const uint16_t* d_q_indices, const float* d_q_values, int n_cand,
                           const int32_t* d_candidates, int64_t n_docs, const int32_t* d_indptr,
                           const uint16_t* d_indices, const float* d_values, int dim, float* d_scores,
                           const float* d_grad_scores, float* d_grad_q, float* d_grad_d,
                           int32_t* d_error, void* stream);

/* write_index(Index*, const char*) -- nsparse/io/index_io.cpp:88-131,170-173.
 * Byte layout identical to the reference writer. */
int nsparse_write_index(nsparse_index* index, const char* path);

/* In-memory writer (BufferedIOWriter, nsparse/io/buffered_io.h:20-62). *out_buf is
 * released with nsparse_buffer_free. */
int nsparse_write_index_buffer(nsparse_index* index, uint8_t** out_buf, size_t* out_size);
void nsparse_buffer_free(uint8_t* buf);

/* read_index(const char*, int flags) -- nsparse/io/index_io.cpp:133-163,175-183.
 * flags & NSPARSE_IO_FLAG_MMAP maps the file while parsing; the index is then
 * uploaded to HBM either way. */
#define NSPARSE_IO_FLAG_MMAP 1 /* IndexIoFlag::kUseMmap, nsparse/io/index_io.h:17-22 */
int nsparse_read_index(const char* path, int flags, nsparse_index** out);
int nsparse_read_index_buffer(const uint8_t* buf, size_t size, int flags, nsparse_index** out);

/* Index::read_csr(path, residency) -- nsparse/index.cpp:63-142 (interchange layout). */
int nsparse_read_csr(nsparse_index* index, const char* path, int residency);

/* Index::get_vectors() (nsparse/index.h:44): copies the stored forward index out.
 * Call with NULL arrays to learn n_rows / nnz / element_size first. Values are the
 * stored bytes (uint8 codes for SQ, float for FP32). */
int nsparse_get_vectors(const nsparse_index* index, int64_t* n_rows, int64_t* nnz, int* element_size,
                        nsparse_idx_t* indptr, nsparse_term_t* indices, uint8_t* values);

#ifdef __cplusplus
}
#endif

#endif /* NSPARSE_C_H */
