// SEISMIC-SQ search of a few queries spread over the whole GPU (uint8 codes,
// block-contiguous rows). The per-query workgroup of search.hip keeps one
// query's traversal on one CU; a single query then reads its rows at a small
// fraction of HBM bandwidth. Here each query takes two launches:
//
//  1. spread_lists_kernel, grid (cut slots, 1 + G, queries): every workgroup
//     loads the query and its cut (top_k_tokens). For cut list ci, workgroup
//     y = 0 computes the list's summary scores (and, for the first list, the
//     visiting order with the step-exact std::sort emulation); workgroups
//     y = 1..G each score the rows of an equal share of the list's docs,
//     writing one key per doc -- (score bits << 32 | ~doc id), 0 for a doc the
//     selector rejects -- and the largest key of each cluster (an atomic max
//     over the shares that hold its docs).
//  2. spread_replay_kernel, one workgroup per query: walks the lists in cut
//     order exactly as the reference does (query_single_inverted_list,
//     seismic_scalar_quantized_index.cpp:67-135; single_query :307-357):
//     prune test against the current k-th score, break on the first list,
//     continue on later ones. A cluster can change the top-k only if it
//     passes its test and its largest key beats the current k-th entry (or
//     the top-k is not full), so 64 clusters are tested at once by one wave,
//     and only those clusters' doc keys are merged, in visiting order; the
//     k-th score moves only at such a merge, so every other cluster's test
//     sees the value it would see in the sequential walk. The top-k equals
//     the reference's under (score desc, doc id asc); a doc seen in an
//     earlier list is either in the top-k (not inserted twice) or worse than
//     its k-th entry, as in the reference, whose visited set skips it.
//
// Every doc of every cut list is scored, so a query reads somewhat more row
// bytes than the pruned walk; the batched search (search.hip) stays the path
// for large batches, where bytes bound the time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

#include "device_index.h"
#include "kernels.h"
#include "search_common.h"
#include "stl_order.h"

namespace nsparse_amd {

namespace {

constexpr int kLBlock = 1024;  // list workgroups (16 waves share the first list's sort)
constexpr int kRBlock = 1024;  // replay workgroup
constexpr int kSubBatch = 2048;  // doc slots per scoring pass of a list workgroup
constexpr int kWinPerStep = 4;   // 64-word windows per wave per step (loads in flight per lane)
constexpr int kMaxCutSlots = 64;
constexpr int kLdsCap = 160 * 1024;

enum ListMisc { kLErr = 0, kLNu, kLNcut, kLFirst, kLNf, kLSortCnt, kLSortCnt1, kLMax, kLMaxCnt, kLMaxAt, kLMiscCount = 16 };
enum ReplayMisc { kRErr = 0, kRTkn, kRMiscCount = 16 };

// LDS and scratch layouts (computed on the host).
// doc keys the replay stages before one merge into the top-k
constexpr int kStage = 64;

struct SpreadPlan {
    int32_t nw;      // query bitmap words
    int32_t qcap;    // longest query (non-zeros) the tables hold
    int32_t mc;      // cluster capacity of a list
    int32_t kmax;    // top-k capacity
    int32_t ncap;    // cut slots (grid.x of the list launch)
    int32_t keycap;  // doc keys of one list the replay stages in LDS
    // list workgroups
    int32_t o_qbits, o_qpref, o_eterm, o_erank, o_eval, o_rlast, o_rval, o_rsum, o_flist, o_cuts, o_wsum, o_misc;
    int32_t o_cstart, o_acc, o_uni, total_list;
    int32_t sort_rng, sort_cap;  // the first-list sort's range lists: byte offset in the union, ranges per level
    // replay workgroup (two buffers per list section)
    int32_t r_cuts, r_misc, r_lmax, r_tks, r_tki, r_ord, r_stg, r_cst, r_ssc, r_cmx, r_key, total_replay;
    int32_t r_cst_stride, r_ssc_stride, r_cmx_stride, r_key_stride;  // bytes between the two buffers
    // per-query scratch, bytes from the query's base
    int64_t s_order, s_sum, s_cmax, s_lmax, s_key, s_bytes;  // s_lmax directly follows the s_cmax section
    int64_t mld;  // doc capacity of a list (row stride of the key section)
};

__device__ __forceinline__ uint64_t doc_key(float s, int32_t id) {
    return (static_cast<uint64_t>(__float_as_uint(s)) << 32) | static_cast<uint32_t>(~id);
}
__device__ __forceinline__ float key_score(uint64_t key) { return __uint_as_float(static_cast<uint32_t>(key >> 32)); }
__device__ __forceinline__ int32_t key_doc(uint64_t key) { return static_cast<int32_t>(~static_cast<uint32_t>(key)); }

// last i in [lo, hi) with a[i] <= v (a[lo] <= v is assumed)
__device__ __forceinline__ int last_le(const int32_t* a, int lo, int hi, int v) {
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Copies n 16-byte units from global to LDS by threads t0, t0 + nt, ...,
// four loads in flight per thread (a thread past the end reloads unit n - 1).
__device__ __forceinline__ void copy_units(uint4* __restrict__ dst, const uint4* __restrict__ src, int n, int t0,
                                           int nt) {
    constexpr int kU = 4;
    for (int i0 = t0; i0 < n; i0 += nt * kU) {
        uint4 v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) v[u] = src[min(i0 + u * nt, n - 1)];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u * nt;
            if (i < n) dst[i] = v[u];
        }
    }
}

// ---------------------------------------------------------------- launch 1

__global__ __launch_bounds__(kLBlock) void spread_lists_kernel(SearchArgs a, SpreadPlan P, unsigned char* scratch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t q = a.query_base + blockIdx.z;
    if (q >= a.n_queries) return;
    const DevIndexView& ix = a.ix;
    const int64_t s0 = a.q_indptr[q];
    const int qn = static_cast<int>(a.q_indptr[q + 1] - s0);
    const int k = a.ks != nullptr ? a.ks[q] : a.k_max;
    const uint32_t* sel = a.sel_all;
    if (a.sel_of_query != nullptr) {
        const int s = a.sel_of_query[q];
        if (s >= a.n_bitmaps) return;  // the replay reports a rejected query
        sel = s >= 0 ? a.sel_bitmaps[s] : nullptr;
    }
    if (qn > P.qcap || k <= 0 || k > P.kmax) return;
    unsigned char* qs = scratch + static_cast<int64_t>(blockIdx.z) * P.s_bytes;

    uint8_t* qcode = smem;
    uint32_t* qbits = reinterpret_cast<uint32_t*>(smem + P.o_qbits);
    uint16_t* qpref = reinterpret_cast<uint16_t*>(smem + P.o_qpref);
    uint16_t* eterm = reinterpret_cast<uint16_t*>(smem + P.o_eterm);
    uint16_t* erank = reinterpret_cast<uint16_t*>(smem + P.o_erank);
    float* evalf = reinterpret_cast<float*>(smem + P.o_eval);
    int32_t* rlast = reinterpret_cast<int32_t*>(smem + P.o_rlast);
    float* rval = reinterpret_cast<float*>(smem + P.o_rval);
    uint32_t* rsum = reinterpret_cast<uint32_t*>(smem + P.o_rsum);
    int32_t* flist = reinterpret_cast<int32_t*>(smem + P.o_flist);
    uint16_t* cuts = reinterpret_cast<uint16_t*>(smem + P.o_cuts);
    int32_t* wsum = reinterpret_cast<int32_t*>(smem + P.o_wsum);
    int32_t* misc = reinterpret_cast<int32_t*>(smem + P.o_misc);
    int32_t* cstart = reinterpret_cast<int32_t*>(smem + P.o_cstart);
    unsigned char* acc = smem + P.o_acc;
    unsigned char* uni = smem + P.o_uni;

    if (tid < kLMiscCount) misc[tid] = 0;
    const QueryTables QT{qbits, qpref, eterm, erank, evalf, rlast, rval, rsum, qcode, wsum, &misc[kLErr], &misc[kLNu]};
    if (!load_query_lds<true, true, kLBlock>(a, s0, qn, P.nw, QT)) return;  // the replay reports it
    // top_k_tokens puts its largest entry first (out of more entries than the
    // cut). With one largest code, whose term has a list, that entry's list is
    // the first the walk visits: its summary workgroup (cut slot 0) starts at
    // once, and another workgroup of slot 0 writes the cut for the replay.
    int fast_term = -1;
    if (blockIdx.x == 0 && blockIdx.y == 0 && qn > a.cut) {
        for (int i = tid; i < qn; i += kLBlock) atomicMax(&misc[kLMax], __float_as_int(evalf[i]));
        __syncthreads();
        const int mx = misc[kLMax];
        for (int i = tid; i < qn; i += kLBlock)
            if (__float_as_int(evalf[i]) == mx) {
                atomicAdd(&misc[kLMaxCnt], 1);
                misc[kLMaxAt] = i;
            }
        __syncthreads();
        if (misc[kLMaxCnt] == 1 && eterm[misc[kLMaxAt]] < static_cast<uint32_t>(ix.n_lists))
            fast_term = eterm[misc[kLMaxAt]];
    }
    if (fast_term < 0 && wave == 0) {
        // top_k_tokens (utils/vector_process.h:18-30) on the codes (cut <= 64)
        const int ncut = topk_tokens_wave(qn, a.cut, evalf, eterm, cuts, lane);
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            int first = -1;  // the first cut list the walk visits (single_query :328-339)
            for (int c = 0; c < ncut && first < 0; ++c)
                if (cuts[c] < static_cast<uint32_t>(ix.n_lists)) first = c;
            misc[kLNcut] = ncut;
            misc[kLFirst] = first;
        }
    }
    __syncthreads();
    const int ncut = fast_term >= 0 ? 1 : misc[kLNcut], first = fast_term >= 0 ? 0 : misc[kLFirst];
    const int ci = blockIdx.x;
    if (ci == 0 && blockIdx.y == 1) {  // the grid has 1 + G >= 2 workgroups per slot
        int32_t* hdr = reinterpret_cast<int32_t*>(qs);
        uint16_t* hcuts = reinterpret_cast<uint16_t*>(qs + 8);
        if (tid == 0) {
            hdr[0] = ncut;
            hdr[1] = first;
        }
        for (int i = tid; i < ncut; i += kLBlock) hcuts[i] = cuts[i];
    }
    if (ci >= ncut) return;
    const uint32_t term = fast_term >= 0 ? static_cast<uint32_t>(fast_term) : cuts[ci];
    if (term >= static_cast<uint32_t>(ix.n_lists)) return;
    const int32_t lb = ix.list_cluster_base[term];
    const int nc = ix.list_cluster_base[term + 1] - lb;
    if (nc == 0 || nc > P.mc) return;

    if (blockIdx.y == 0) {
        // ---- summary scores of list ci (and the first list's visiting order) ----
        const int32_t tb = ix.list_term_base[term], te = ix.list_term_base[term + 1];
        int32_t* rcol = rlast;  // free after the query setup
        uint32_t* sacc = reinterpret_cast<uint32_t*>(acc);
        const int nu = misc[kLNu];
        for (int r = tid; r < nu; r += kLBlock) rcol[r] = -1;
        for (int c = tid; c < nc; c += kLBlock) sacc[c] = 0;
        __syncthreads();
        const QueryLds QL{qbits, qpref, rval};
        match_summary_columns<kLBlock>(ix, term, tb, te, qn, eterm, erank, QL, rcol);
        __syncthreads();
        for (int r = tid; r < nu; r += kLBlock)
            if (rcol[r] >= 0) flist[atomicAdd(&misc[kLNf], 1)] = r;
        __syncthreads();
        accumulate_summaries_sq<kLBlock>(ix, tb, te, misc[kLNf], flist, rcol, rsum, uni, sacc, P.qcap);
        uint32_t* gsum = reinterpret_cast<uint32_t*>(qs + P.s_sum) + static_cast<int64_t>(ci) * P.mc;
        for (int c = tid; c < nc; c += kLBlock) gsum[c] = sacc[c];
        if (ci == first) {
            // reorder_clusters (seismic_common.h:105-115)
            KeyItem* A = reinterpret_cast<KeyItem*>(uni);
            for (int c = tid; c < nc; c += kLBlock) A[c] = KeyItem{static_cast<float>(sacc[c]), static_cast<uint32_t>(c)};
            __syncthreads();
            uint32_t* leaf_bits = reinterpret_cast<uint32_t*>(flist);
            uint16_t* Lpos = reinterpret_cast<uint16_t*>(A + P.mc);
            introsort_desc_block<kLBlock>(A, nc, Lpos, Lpos + P.mc, leaf_bits,
                                          reinterpret_cast<int4*>(uni + P.sort_rng), P.sort_cap, &misc[kLSortCnt]);
            introsort_final_pass(A, nc, leaf_bits, reinterpret_cast<uint16_t*>(qs + P.s_order), tid, kLBlock);
        }
        return;
    }

    // ---- doc keys of a slice of list ci's docs ----
    const int G = static_cast<int>(gridDim.y) - 1, g = static_cast<int>(blockIdx.y) - 1;
    const int64_t dbase = ix.cluster_doc_ptr[lb];
    for (int j = tid; j <= nc; j += kLBlock) cstart[j] = static_cast<int32_t>(ix.cluster_doc_ptr[lb + j] - dbase);
    __syncthreads();
    // docs [p0, p1): an equal share of the list (a cluster can span shares);
    // clusters [c0, c1) hold them
    const int nd_list = cstart[nc];
    const int p0 = static_cast<int>(static_cast<int64_t>(nd_list) * g / G);
    const int p1 = static_cast<int>(static_cast<int64_t>(nd_list) * (g + 1) / G);
    if (p0 >= p1) return;
    const int c0 = last_le(cstart, 0, nc, p0);
    const int c1 = last_le(cstart, 0, nc, p1 - 1) + 1;
    unsigned long long* lcmax = reinterpret_cast<unsigned long long*>(acc);
    for (int c = c0 + tid; c < c1; c += kLBlock) lcmax[c - c0] = 0;
    int32_t* wst = reinterpret_cast<int32_t*>(uni);                // kSubBatch + 1 word starts
    uint32_t* isc = reinterpret_cast<uint32_t*>(wst + kSubBatch + 1);  // kSubBatch sums
    uint64_t* gkey = reinterpret_cast<uint64_t*>(qs + P.s_key) + static_cast<int64_t>(ci) * P.mld;
    const uint4* rows = ix.blk_rows;
    for (int pa = p0; pa < p1; pa += kSubBatch) {
        const int nd = min(p1 - pa, kSubBatch);
        __syncthreads();  // the previous pass is done with wst / isc
        // the rows of consecutive postings are consecutive words
        const int64_t w0 = blk_ref_first(ix.row_ref[dbase + pa]);
        for (int j = tid; j < nd; j += kLBlock) {
            const int64_t ref = ix.row_ref[dbase + pa + j];
            const int ws = static_cast<int>(blk_ref_first(ref) - w0);
            wst[j] = ws;
            isc[j] = 0;
            if (j == nd - 1) wst[nd] = ws + blk_ref_words(ref);
        }
        __syncthreads();
        // Windows of 64 words, one per lane; a DPP prefix sum S over the window
        // gives each doc run's sum: the last lane of a run adds S to its doc
        // and subtracts S from the next run's doc (integer LDS atomics, exact
        // in any order). Rows are zero padded to whole words; lanes past the
        // end read the zero word after the last row.
        const int W = wst[nd];
        const int nwin = (W + 63) >> 6;
        for (int x0 = wave * kWinPerStep; x0 < nwin; x0 += (kLBlock / 64) * kWinPerStep) {
            int jv[kWinPerStep];
            int64_t wv[kWinPerStep];
#pragma unroll
            for (int u = 0; u < kWinPerStep; ++u) {
                const int f = ((x0 + u) << 6) + lane;
                const bool has = f < W;
                jv[u] = has ? last_le(wst, 0, nd, f) : -1;
                wv[u] = has ? w0 + f : ix.blk_words;
            }
            uint4 ii[kWinPerStep];
#pragma unroll
            for (int u = 0; u < kWinPerStep; ++u) ii[u] = rows[wv[u]];
            uint32_t sums[kWinPerStep];
#pragma unroll
            for (int u = 0; u < kWinPerStep; ++u) {
                const uint32_t s = blk_word_dot(ii[u], qcode, 0U);
                sums[u] = s;
            }
#pragma unroll
            for (int u = 0; u < kWinPerStep; ++u) sums[u] = wave_incl_scan_dpp(sums[u]);
#pragma unroll
            for (int u = 0; u < kWinPerStep; ++u) {
                const int f = ((x0 + u) << 6) + lane;
                const int j = jv[u];
                // the doc of the next lane's word (a DPP shift left by one lane)
                const int j1 = __builtin_amdgcn_update_dpp(0, j, 0x130, 0xF, 0xF, false);
                const uint32_t S = sums[u];
                if (j >= 0 && S != 0 && (j1 != j || lane == 63 || f + 1 >= W)) {
                    atomicAdd(&isc[j], S);
                    if (lane != 63 && f + 1 < W && j1 != j) atomicSub(&isc[j1], S);
                }
            }
        }
        __syncthreads();
        for (int j = tid; j < nd; j += kLBlock) {
            const int p = pa + j;
            const int32_t id = ix.docs[dbase + p];
            const bool keep = sel == nullptr || ((sel[id >> 5] >> (id & 31)) & 1u);
            const uint64_t key = keep ? doc_key(static_cast<float>(isc[j]), id) : 0;
            gkey[p] = key;
            atomicMax(&lcmax[last_le(cstart, c0, c1, p) - c0], static_cast<unsigned long long>(key));
        }
    }
    __syncthreads();
    // the largest key of each cluster, over the shares that hold its docs,
    // and of the list (zeroed before the launch)
    unsigned long long* gcmax = reinterpret_cast<unsigned long long*>(qs + P.s_cmax) + static_cast<int64_t>(ci) * P.mc;
    unsigned long long lmx = 0;
    for (int c = c0 + tid; c < c1; c += kLBlock) {
        const unsigned long long v = lcmax[c - c0];
        if (v != 0) atomicMax(&gcmax[c], v);
        lmx = v > lmx ? v : lmx;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const uint32_t hi = __shfl_xor(static_cast<uint32_t>(lmx >> 32), d, 64);
        const uint32_t lo = __shfl_xor(static_cast<uint32_t>(lmx), d, 64);
        const unsigned long long o = (static_cast<unsigned long long>(hi) << 32) | lo;
        lmx = o > lmx ? o : lmx;
    }
    if (lane == 0 && lmx != 0)
        atomicMax(reinterpret_cast<unsigned long long*>(qs + P.s_lmax) + ci, lmx);
}

// ---------------------------------------------------------------- launch 2

__global__ __launch_bounds__(kRBlock) void spread_replay_kernel(SearchArgs a, SpreadPlan P,
                                                              const unsigned char* scratch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t q = a.query_base + blockIdx.x;
    if (q >= a.n_queries) return;
    const DevIndexView& ix = a.ix;
    const int64_t s0 = a.q_indptr[q];
    const int qn = static_cast<int>(a.q_indptr[q + 1] - s0);
    const int k = a.ks != nullptr ? a.ks[q] : a.k_max;
    float* out_d = a.out_d + q * a.k_max;
    int32_t* out_l = a.out_l + q * a.k_max;
    bool bad_sel = false;
    if (a.sel_of_query != nullptr) bad_sel = a.sel_of_query[q] >= a.n_bitmaps;
    int32_t* misc = reinterpret_cast<int32_t*>(smem + P.r_misc);
    if (tid < kRMiscCount) misc[tid] = 0;
    __syncthreads();
    int err = bad_sel ? kDevBadSelector : qn > P.qcap ? kDevQueryTooLong : (k <= 0 || k > P.kmax) ? kDevBadK : 0;
    if (err == 0) {
        for (int i = tid; i < qn; i += kRBlock)
            if (a.q_indices[s0 + i] >= static_cast<uint32_t>(ix.dim)) misc[kRErr] = 1;
        __syncthreads();
        if (misc[kRErr] != 0) err = kDevTermOutOfRange;
    }
    if (err != 0) {
        if (tid == 0) atomicCAS(a.error_flag, 0, err);
        for (int i = tid; i < a.k_max; i += kRBlock) {
            out_d[i] = -1.0F;
            out_l[i] = -1;
        }
        return;
    }
    const unsigned char* qs = scratch + static_cast<int64_t>(blockIdx.x) * P.s_bytes;
    const int ncut = reinterpret_cast<const int32_t*>(qs)[0];
    const int first_ci = reinterpret_cast<const int32_t*>(qs)[1];
    uint16_t* cuts = reinterpret_cast<uint16_t*>(smem + P.r_cuts);
    float* tks = reinterpret_cast<float*>(smem + P.r_tks);
    int32_t* tki = reinterpret_cast<int32_t*>(smem + P.r_tki);
    uint16_t* ord = reinterpret_cast<uint16_t*>(smem + P.r_ord);
    uint64_t* lmax = reinterpret_cast<uint64_t*>(smem + P.r_lmax);
    // snap[j & 1]: the k-th entry's key before list j is walked, 0 while the
    // top-k is not full
    uint64_t* snap = reinterpret_cast<uint64_t*>(misc + 4);
    for (int i = tid; i < ncut; i += kRBlock) {
        cuts[i] = reinterpret_cast<const uint16_t*>(qs + 8)[i];
        lmax[i] = reinterpret_cast<const uint64_t*>(qs + P.s_lmax)[i];
    }
    if (tid < 2) snap[tid] = 0;
    __syncthreads();
    const float hf = a.heap_factor;

    // the list sections of buffer b
    auto cst_of = [&](int b) { return reinterpret_cast<int32_t*>(smem + P.r_cst + b * P.r_cst_stride); };
    auto ssc_of = [&](int b) { return reinterpret_cast<float*>(smem + P.r_ssc + b * P.r_ssc_stride); };
    auto cmx_of = [&](int b) { return reinterpret_cast<uint64_t*>(smem + P.r_cmx + b * P.r_cmx_stride); };
    auto key_of = [&](int b) { return reinterpret_cast<uint64_t*>(smem + P.r_key + b * P.r_key_stride); };
    const uint32_t* gsum_all = reinterpret_cast<const uint32_t*>(qs + P.s_sum);
    const uint64_t* gcmax_all = reinterpret_cast<const uint64_t*>(qs + P.s_cmax);
    const uint64_t* gkey_all = reinterpret_cast<const uint64_t*>(qs + P.s_key);

    // Stages list ci's cluster starts, summary scores, largest keys, doc keys
    // (when they fit) and, for the first list, its visiting order into buffer
    // b, by threads t0, t0 + nt, ...
    auto stage = [&](int ci, int b, int t0, int nt) {
        const uint32_t term = cuts[ci];
        if (term >= static_cast<uint32_t>(ix.n_lists)) return;
        const int32_t lb = ix.list_cluster_base[term];
        const int nc = ix.list_cluster_base[term + 1] - lb;
        if (nc == 0 || nc > P.mc) return;
        const int64_t dbase = ix.cluster_doc_ptr[lb];
        const int64_t nd = ix.cluster_doc_ptr[lb + nc] - dbase;
        int32_t* cst = cst_of(b);
        float* ssc = ssc_of(b);
        for (int c = t0; c <= nc; c += nt) cst[c] = static_cast<int32_t>(ix.cluster_doc_ptr[lb + c] - dbase);
        const uint32_t* gs = gsum_all + static_cast<int64_t>(ci) * P.mc;
        for (int c = t0; c < nc; c += nt) ssc[c] = static_cast<float>(gs[c]);
        // 16-byte units: P.mc and P.mld are even, the sections 16-byte aligned
        copy_units(reinterpret_cast<uint4*>(cmx_of(b)),
                   reinterpret_cast<const uint4*>(gcmax_all + static_cast<int64_t>(ci) * P.mc), (nc + 1) >> 1, t0,
                   nt);
        if (nd <= P.keycap)
            copy_units(reinterpret_cast<uint4*>(key_of(b)),
                       reinterpret_cast<const uint4*>(gkey_all + static_cast<int64_t>(ci) * P.mld),
                       static_cast<int>((nd + 1) >> 1), t0, nt);
        if (ci == first_ci) {
            const uint16_t* go = reinterpret_cast<const uint16_t*>(qs + P.s_order);
            for (int c = t0; c < nc; c += nt) ord[c] = go[c];
        }
    };

    // The exact walk of list ci from buffer b (wave 0). Doc keys that beat the
    // k-th entry are staged (up to kStage) and merged into the top-k together:
    // the top-k of a set of distinct keys does not depend on the order they
    // arrive in, so a batch equals the one-by-one merges as long as the walk
    // visits the same clusters. A test that fails against the merged top-k
    // fails in the sequential walk too (its k-th score only rises). A test
    // passes for sure if it passes against the highest k-th score the staged
    // keys can lead to: with sm of them, the entry now at k - 1 - sm (a
    // non-full top-k stays below k while tkn + sm < k). Where that is not
    // certain the staged keys are merged first and the test runs exactly.
    // Staging a doc that in the sequential walk would not beat the k-th entry
    // changes nothing: it stays below every key of the final top-k.
    int tkn = 0;
    int sm = 0;  // staged keys
    uint64_t* stg = reinterpret_cast<uint64_t*>(smem + P.r_stg);
    auto flush = [&] {
        if (sm == 0) return;
        __builtin_amdgcn_wave_barrier();
        const uint64_t kv = lane < sm ? stg[lane] : 0;
        __builtin_amdgcn_wave_barrier();
        merge_topk_wave<true>(kv != 0, key_score(kv), key_doc(kv), k, tkn, tks, tki, lane);
        sm = 0;
    };
    auto replay = [&](int ci, int b) {
        const uint32_t term = cuts[ci];
        if (term >= static_cast<uint32_t>(ix.n_lists)) return;
        const int32_t lb = ix.list_cluster_base[term];
        const int nc = ix.list_cluster_base[term + 1] - lb;
        if (nc == 0 || nc > P.mc) return;
        const int64_t nd = ix.cluster_doc_ptr[lb + nc] - ix.cluster_doc_ptr[lb];
        const bool first = ci == first_ci;
        const int32_t* cst = cst_of(b);
        const float* ssc = ssc_of(b);
        const uint64_t* cmx = cmx_of(b);
        const uint64_t* keys = nd <= P.keycap ? key_of(b) : gkey_all + static_cast<int64_t>(ci) * P.mld;
        for (int pos = 0; pos < nc; pos += 64) {
            const int p = pos + lane;
            const bool valid = p < nc;
            const int c = valid ? (first ? static_cast<int>(ord[p]) : p) : 0;
            const float s = valid ? ssc[c] : 0.0F;
            const uint64_t mx = valid ? cmx[c] : 0;
            uint64_t done = 0;  // lanes (visiting positions) already walked
            bool brk = false;
            while (true) {
                // tests against the merged top-k (certain fails; passes are
                // checked against the staged keys below)
                const bool full = tkn == k;
                const float kth = full ? tks[k - 1] : 0.0F;
                const uint64_t tkey = full ? doc_key(kth, tki[k - 1]) : 0;
                const bool fail = valid && full && (s * hf < kth);
                const bool live = valid && !fail && mx != 0 && (!full || mx > tkey);
                const uint64_t mf = ballot(fail) & ~done;
                uint64_t m = ballot(live) & ~done;
                if (first && mf != 0) {
                    // the first list stops at its first failed test
                    m &= (1ULL << (__ffsll(static_cast<unsigned long long>(mf)) - 1)) - 1;
                    if (m == 0) {
                        brk = true;
                        break;
                    }
                }
                if (m == 0) break;
                const int i = __ffsll(static_cast<unsigned long long>(m)) - 1;
                const uint64_t upto = (2ULL << i) - 1;  // lanes <= i (all 64 when i = 63)
                if (sm > 0) {
                    // lane i's pass, and on the first list every pass before
                    // it (a failure there would end the walk), must hold
                    // whatever the staged keys do to the k-th score
                    bool sure;
                    if (!full) {
                        sure = tkn + sm < k;
                    } else if (sm >= k) {
                        sure = false;
                    } else {
                        const float bound = tks[k - 1 - sm];
                        const uint64_t scope = first ? upto : (1ULL << i);
                        sure = (ballot(valid && !fail && s * hf < bound) & ~done & scope) == 0;
                    }
                    if (!sure) {
                        flush();
                        continue;
                    }
                }
                const int cc = __builtin_amdgcn_readlane(c, i);
                const int d1 = cst[cc + 1];
                // The cluster's docs, 64 at a time with kU loads in flight;
                // those that beat the merged top-k's k-th entry are staged.
                constexpr int kU = 4;
                uint64_t kk = tkey;
                for (int d = cst[cc]; d < d1; d += 64 * kU) {
                    uint64_t kv[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int dd = d + u * 64 + lane;
                        kv[u] = dd < d1 ? keys[dd] : 0;
                    }
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        bool take = kv[u] != 0 && (tkn < k || kv[u] > kk);
                        uint64_t tm = ballot(take);
                        if (tm == 0) continue;
                        if (sm + __popcll(tm) > kStage) {
                            flush();
                            kk = tkn == k ? doc_key(tks[k - 1], tki[k - 1]) : 0;
                            take = kv[u] != 0 && (tkn < k || kv[u] > kk);
                            tm = ballot(take);
                            if (tm == 0) continue;
                        }
                        if (take) stg[sm + __popcll(tm & ((1ULL << lane) - 1))] = kv[u];
                        sm += __popcll(tm);
                    }
                }
                done |= upto;
                if (done == ~0ULL) break;
            }
            // the first list's next window (or its end) must see every test
            // exact: a failure among this window's unchecked passes ends it
            if (first) flush();
            if (brk) break;
        }
        flush();
    };

    // List ci is walked from buffer ci & 1 while the other waves stage list
    // ci + 1 into the other buffer. Once the top-k is full, a list whose
    // largest key does not beat its k-th entry can change nothing: it is
    // neither staged nor walked. The stagers decide on the k-th entry before
    // list ci, the walk on the current one (never lower), so a list that is
    // walked has been staged.
    if (ncut > 0) stage(0, 0, tid, kRBlock);
    __syncthreads();
    for (int ci = 0; ci < ncut; ++ci) {
        if (wave == 0) {
            const uint64_t now = tkn == k ? doc_key(tks[k - 1], tki[k - 1]) : 0;
            if (now == 0 || lmax[ci] > now) replay(ci, ci & 1);
            const uint64_t after = tkn == k ? doc_key(tks[k - 1], tki[k - 1]) : 0;
            if (lane == 0) snap[(ci + 1) & 1] = after;
        } else if (ci + 1 < ncut) {
            const uint64_t before = snap[ci & 1];
            if (before == 0 || lmax[ci + 1] > before) stage(ci + 1, (ci + 1) & 1, tid - 64, kRBlock - 64);
        }
        __syncthreads();
    }
    if (tid == 0) misc[kRTkn] = tkn;
    __syncthreads();
    const int n_out = misc[kRTkn];
    for (int i = tid; i < a.k_max; i += kRBlock) {
        float d = -1.0F;
        int32_t l = -1;
        if (i < k && i < n_out) {
            d = tks[i] * a.decode_scale;
            l = tki[i];
        }
        out_d[i] = d;
        out_l[i] = l;
    }
}

// Clears the error flag and every query's largest-key sections.
__global__ void spread_prepare_kernel(int32_t* flag, unsigned char* scratch, int64_t s_bytes, int64_t s_cmax,
                                      int64_t words, int64_t n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
    const int64_t total = words * n;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        reinterpret_cast<uint64_t*>(scratch + (i / words) * s_bytes + s_cmax)[i % words] = 0;
}

__global__ void list_docs_max_kernel(const int32_t* __restrict__ lcb, const int64_t* __restrict__ cdp, int n_lists,
                                     unsigned long long* __restrict__ out) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < n_lists; l += gridDim.x * blockDim.x)
        atomicMax(out, static_cast<unsigned long long>(cdp[lcb[l + 1]] - cdp[lcb[l]]));
}

int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

SpreadPlan make_spread_plan(int dim, int max_clusters, int kmax, int cut, int64_t max_list_docs, int qcap) {
    SpreadPlan P{};
    P.nw = (dim + 31) / 32;
    P.qcap = std::max(kQMax, (qcap + 63) & ~63);
    P.mc = std::max(2, (max_clusters + 1) & ~1);
    P.kmax = kmax;
    P.ncap = std::max(1, std::min(cut, P.qcap));
    P.mld = std::max<int64_t>(2, (max_list_docs + 1) & ~int64_t{1});
    int o = 0;
    auto take = [&](int bytes) {
        const int at = o;
        o += (bytes + 15) & ~15;
        return at;
    };
    take(dim);  // the dense code table at offset 0
    P.o_qbits = take(P.nw * 4);
    P.o_qpref = take(P.nw * 2);
    P.o_eterm = take(P.qcap * 2);
    P.o_erank = take(P.qcap * 2);
    P.o_eval = take(P.qcap * 4);
    P.o_rlast = take(P.qcap * 4);
    P.o_rval = take(P.qcap * 4);
    P.o_rsum = take(P.qcap * 4);
    P.o_flist = take(P.qcap * 4);
    P.o_cuts = take(P.qcap * 2);
    P.o_wsum = take(kMaxWaves * 4);
    P.o_misc = take(kLMiscCount * 4);
    P.o_cstart = take((P.mc + 1) * 4);
    P.o_acc = take(P.mc * 8);
    // summary scratch, the cut heap, the sort buffers, or one scoring pass
    P.sort_cap = 2 * P.mc / 17 + 2;
    P.sort_rng = (P.mc * 12 + 15) & ~15;
    const int sort_bytes = P.sort_rng + 2 * P.sort_cap * 16;
    const int uni = std::max(std::max(P.qcap * 12 + 16, P.qcap * 8), std::max(sort_bytes, (2 * kSubBatch + 1) * 4));
    P.o_uni = take(uni);
    P.total_list = o;

    o = 0;
    P.r_cuts = take(P.qcap * 2);
    P.r_misc = take(kRMiscCount * 4);
    P.r_lmax = take(P.ncap * 8);
    P.r_tks = take(kmax * 4);
    P.r_tki = take(kmax * 4);
    P.r_ord = take(P.mc * 2);
    P.r_stg = take(kStage * 8);
    P.r_cst_stride = ((P.mc + 1) * 4 + 15) & ~15;
    P.r_ssc_stride = (P.mc * 4 + 15) & ~15;
    P.r_cmx_stride = P.mc * 8;
    P.r_cst = take(2 * P.r_cst_stride);
    P.r_ssc = take(2 * P.r_ssc_stride);
    P.r_cmx = take(2 * P.r_cmx_stride);
    // doc keys: as many as fit, up to the longest list
    const int64_t room = (kLdsCap - o) / 16;  // per buffer, in keys
    P.keycap = static_cast<int32_t>(std::max<int64_t>(0, std::min<int64_t>(room & ~int64_t{1}, P.mld)));
    P.r_key_stride = P.keycap * 8;
    P.r_key = take(2 * P.r_key_stride);
    P.total_replay = o;

    int64_t s = align_up(8 + P.qcap * 2, 256);
    P.s_order = s;
    s = align_up(s + P.mc * 2, 256);
    P.s_sum = s;
    s = align_up(s + static_cast<int64_t>(P.ncap) * P.mc * 4, 256);
    P.s_cmax = s;
    P.s_lmax = s + static_cast<int64_t>(P.ncap) * P.mc * 8;
    s = align_up(P.s_lmax + static_cast<int64_t>(P.ncap) * 8, 256);
    P.s_key = s;
    s = align_up(s + static_cast<int64_t>(P.ncap) * P.mld * 8, 256);
    P.s_bytes = s;
    return P;
}

int cu_count() {
    static const int cus = [] {
        int dev = 0, n = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        return n;
    }();
    return cus;
}

}  // namespace

bool search_spread_fits(const DevIndexView& ix, int max_clusters, int kmax, int cut, int64_t max_list_docs,
                        int qcap) {
    if (ix.esz != 1 || ix.row_ref == nullptr || cut <= 0 || std::min(cut, std::max(qcap, kQMax)) > kMaxCutSlots)
        return false;
    // cluster ids are 16-bit in the visiting order; doc positions 32-bit
    if (max_clusters <= 0 || max_clusters > 16384 || max_list_docs >= (int64_t{1} << 31)) return false;
    const SpreadPlan P = make_spread_plan(ix.dim, max_clusters, kmax, cut, max_list_docs, qcap);
    return P.total_list <= kLdsCap && P.total_replay <= kLdsCap && P.keycap > 0;
}

size_t search_spread_scratch_bytes(const DevIndexView& ix, int max_clusters, int kmax, int cut,
                                   int64_t max_list_docs, int64_t n_queries, int qcap) {
    const SpreadPlan P = make_spread_plan(ix.dim, max_clusters, kmax, cut, max_list_docs, qcap);
    return static_cast<size_t>(P.s_bytes) * static_cast<size_t>(n_queries);
}

void launch_search_spread(const SearchArgs& args, int max_clusters, int kmax, int64_t max_list_docs, void* scratch,
                          hipStream_t stream, int qcap) {
    const SpreadPlan P = make_spread_plan(args.ix.dim, max_clusters, kmax, args.cut, max_list_docs, qcap);
    if (!search_spread_fits(args.ix, max_clusters, kmax, args.cut, max_list_docs, qcap))
        throw std::invalid_argument("the spread search does not fit this index");
    if (args.error_flag == nullptr || scratch == nullptr)
        throw std::invalid_argument("spread search launch without an error flag or scratch");
    if (args.stats != nullptr) throw std::invalid_argument("the spread search keeps no traversal counters");
    if (args.n_queries <= 0 || args.n_queries > 65535) throw std::invalid_argument("spread search batch size");
    // doc-scoring workgroups per cut list: one wave of list workgroups fills the GPU
    int per_cu = 1;
    {
        static std::mutex mu;
        static bool attrs = false;
        static int cached = 0, cached_lds = -1;
        std::lock_guard<std::mutex> lock(mu);
        if (!attrs) {
            hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(spread_lists_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, kLdsCap),
                      "hipFuncSetAttribute");
            hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(spread_replay_kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, kLdsCap),
                      "hipFuncSetAttribute");
            attrs = true;
        }
        if (cached_lds != P.total_list) {
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &cached, reinterpret_cast<const void*>(spread_lists_kernel), kLBlock, P.total_list),
                      "occupancy");
            cached_lds = P.total_list;
        }
        per_cu = std::max(cached, 1);
    }
    // one query's lists fill the GPU; a launch of several shares one fill among
    // them (at least 4 shares per list), since the per-workgroup query load and
    // cut selection, not the row bytes, set the time of small shares
    const int64_t fill = static_cast<int64_t>(per_cu) * cu_count();
    int groups = static_cast<int>(std::max<int64_t>(1, fill / P.ncap - 1));
    if (args.n_queries > 1)
        groups = static_cast<int>(std::max<int64_t>(std::min(groups, 4), fill / (P.ncap * args.n_queries) - 1));
    if (const char* e = std::getenv("NSPARSE_SPREAD_GROUPS")) groups = std::max(1, std::atoi(e));
    groups = std::min(groups, 1024);
    SearchArgs a = args;
    a.query_base = 0;
    auto* s = static_cast<unsigned char*>(scratch);
    const int64_t words = static_cast<int64_t>(P.ncap) * (P.mc + 1);  // the cluster and list maxima
    const unsigned zgrid = static_cast<unsigned>(std::min<int64_t>((words * a.n_queries + 255) / 256, 1024));
    hipLaunchKernelGGL(spread_prepare_kernel, dim3(zgrid), dim3(256), 0, stream, a.error_flag, s, P.s_bytes,
                       P.s_cmax, words, a.n_queries);
    hipLaunchKernelGGL(spread_lists_kernel, dim3(static_cast<unsigned>(P.ncap), static_cast<unsigned>(groups + 1),
                                               static_cast<unsigned>(a.n_queries)),
                       dim3(kLBlock), P.total_list, stream, a, P, s);
    hipLaunchKernelGGL(spread_replay_kernel, dim3(static_cast<unsigned>(a.n_queries)), dim3(kRBlock), P.total_replay,
                       stream, a, P, static_cast<const unsigned char*>(s));
    hip_check(hipGetLastError(), "spread search launch");
}

int64_t max_list_docs(const DevIndexView& ix) {
    if (ix.n_lists <= 0 || ix.list_cluster_base == nullptr || ix.cluster_doc_ptr == nullptr) return 0;
    DeviceBuffer out(8);
    hip_check(hipMemset(out.get(), 0, 8), "hipMemset");
    const unsigned grid = static_cast<unsigned>(std::min((ix.n_lists + 255) / 256, 1024));
    hipLaunchKernelGGL(list_docs_max_kernel, dim3(grid), dim3(256), 0, nullptr, ix.list_cluster_base,
                       ix.cluster_doc_ptr, ix.n_lists, out.get<unsigned long long>());
    hip_check(hipGetLastError(), "list size launch");
    int64_t v = 0;
    out.download(&v, 8);
    return v;
}

}  // namespace nsparse_amd
