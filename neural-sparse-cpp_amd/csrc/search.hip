// SEISMIC / SEISMIC-SQ search on CDNA4 (gfx950).
//
// One 256-lane workgroup per query. The query lives in LDS as a presence
// bitmap over the vocabulary plus a rank table (popcount prefix per 32-term
// word) into compact per-term values -- about 6 KB for a 30522-term vocabulary
// instead of a 30-122 KB dense vector -- so several queries share a CU.
//
// Per cut list the workgroup (reference: query_single_inverted_list,
// seismic_scalar_quantized_index.cpp:67-135 and seismic_index.cpp:53-118):
//  1. scans the list's summary term ids against the query bitmap and
//     accumulates block-summary scores from the matching CSC columns
//     (integer-exact LDS atomics for uint8 codes; for FP32 a term-ordered tile
//     so every cluster adds its products in query-entry order, as the
//     reference's score_summaries_typed does);
//  2. orders clusters (first list only) with a step-exact std::sort emulation;
//  3. walks clusters in that order in speculative chunks: wave 0 picks the
//     next run of clusters that pass the heap-factor test against the current
//     k-th score (a cluster failing an earlier, lower k-th score also fails
//     every later one), all lanes gather those clusters' doc rows from HBM and
//     score them, then wave 0 replays the chunk in order against the exact,
//     evolving k-th score -- so every prune decision equals the reference's.
// The top-k is a sorted LDS array ordered by (score desc, doc id asc); a doc
// already in it is not inserted twice, which reproduces the reference's
// visited-set semantics for the result (a revisited doc has the same score).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <cstdlib>

#include <algorithm>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>

#include "device_index.h"
#include "kernels.h"
#include "search_common.h"
#include "stl_order.h"

namespace nsparse_amd {

constexpr int kBlock = 256;  // exact-match, brute-force and default search workgroups
constexpr int kWaves = kBlock / 64;
// Doc slots per speculative chunk. Two chunks are staged at once (one is
// scored while the next is gathered), 20 B per slot each plus a word-start
// table: 512 slots with the dense code table keep its LDS plan within two
// 512-lane workgroups per CU, 256 with the bitmap lookup or FP32 rows stay
// within four workgroups per CU. A 1024-lane workgroup (alone on its CU) could
// stage 2048, but smaller chunks speculate less past the prune point and turn
// over sooner: one query per launch at 8.8M docs took p50 1.27 ms with 2048,
// 1.17 ms with 1024, 1.14 ms with 768 or 512; 384 was 2% slower there and
// 1.1% slower for the batched 512-lane launch. A multiple of 64 up to 4096
// (the first-doc search reads every budget/64-th word start with one lane each).
// The first-list sort's range lists (introsort_desc_block) in the union
// scratch, after KeyItem[mc] and two uint16[mc] position lists: two levels of
// at most 2 * mc / 17 + 1 ranges each.
__host__ __device__ constexpr int sort_range_cap(int mc) { return 2 * mc / 17 + 2; }
__host__ __device__ constexpr int sort_rng_offset(int mc) { return (mc * 12 + 15) & ~15; }

__host__ __device__ constexpr int chunk_budget(bool dense, int /*block*/) {
    return dense ? 512 : 256;
}
constexpr int kSegMax = 64;   // clusters examined per chunk (one wave pass)
constexpr int kTileFloats = 4096;
// windows per wave per scoring step (loads in flight per lane) for
// forward-index rows: 5 measured best at 8.8M docs with 24-byte row words
// (3: -4%, 4: -1.1%, 6: -0.1%); the kernel stays within 128 VGPRs, i.e. four
// waves per SIMD
template <bool DENSE>
constexpr int win_per_step() {
    return 5;
}
// the same for block rows, whose lanes each hold a 32-byte pair of words: 4
// measured best at 8.8M docs (29.04-29.10 ms per 10k-query launch; 3:
// 29.45 ms, 5: 29.12-29.18 ms)
constexpr int kBlkWinPerStep = 4;

// A piece of one cluster's docs in a chunk. Cluster ids and positions fit 16
// bits: the LDS plan holds 10 bytes per cluster, so a list has < 16384.
struct Seg {
    int32_t doc_begin;  // offset into the list's docs
    uint16_t cluster;   // local cluster id
    uint16_t pos;       // position of the cluster in the visiting order
    uint16_t count;     // docs in this piece
    uint16_t slot;      // first staging slot
    uint16_t test;      // 1 if this piece starts the cluster (prune test runs here)
    uint16_t pad;
};

// One staged chunk: per doc slot the doc id, the row start (block rows: first
// word minus the doc's first chunk word), the row length (-1 when the
// selector rejects the doc) and the score (uint32 sum for uint8 codes, float
// for FP32); the word-start table (budget + 128 entries); the pieces; and
// meta[kChunkMetaCount].
struct Chunk {
    int32_t* id;
    int32_t* len;
    int32_t* rs;
    uint32_t* isc;
    float* sc;
    int32_t* wbase;
    Seg* seg;
    int32_t* meta;
};
enum ChunkMeta {
    kChunkNseg = 0,  // pieces
    kChunkTotal,     // staged doc slots
    kChunkWords,     // 8-element words of the staged rows
    kChunkNext,      // next window (or 64-slot group) to score
    kChunkEnd,       // 1: no chunk here, the list's traversal is over
    kChunkStamp,     // index of the chunk this slot holds, written last (release)
    kChunkCheckin,   // waves that finished scoring the chunk
    kChunkMetaCount
};

__device__ __forceinline__ int lds_load_acquire(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_store_release(int32_t* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

enum Misc {
    kMiscErr = 0,
    kMiscNu,     // unique query terms
    kMiscNcut,   // cut terms
    kMiscNf,     // found items in the current list
    kMiscPos,    // next cluster position
    kMiscPoff,   // docs of the cluster at kMiscPos already handed out
    kMiscEnd,    // clusters at positions >= this are never visited (first list break)
    kMiscDone,   // list finished
    kMiscTkN,    // entries in the top-k
    kMiscBreakPending,
    kMiscStatFound,  // counters of the current list (stats runs only)
    kMiscStatCsc,
    kMiscSortCnt,    // two ints: the first-list sort's range counts
    kMiscCount = 16
};

// Byte offsets of the dynamic LDS sections (computed on the host).
struct LdsPlan {
    int32_t nw;        // 32-bit words of the query bitmap
    int32_t budget;    // doc slots per speculative chunk
    int32_t mc;        // cluster capacity
    int32_t kmax;      // top-k capacity
    int32_t uni_bytes; // union scratch (staging / sort buffer / FP32 tile)
    int32_t o_qbits, o_qpref, o_eterm, o_erank, o_eval, o_rlast, o_rval, o_rsum, o_rcol, o_flist, o_cuts;
    int32_t o_cstart, o_score, o_order, o_uni, o_tks, o_tki, o_seg, o_misc, o_wsum, o_queue, o_chunk, o_qstat;
    int32_t total;
};

// DENSE (uint8 codes only): the query's codes also sit in a dense
// vocabulary-sized byte table at LDS offset 0, so the doc-row scoring loop
// does one byte read per element instead of bitmap + rank + value reads; it
// costs dim bytes of LDS per workgroup.
// W16 (with SQ false): 16-bit codes (QuantizerType::QT_16bit). The traversal
// is the FP32 one (bitmap + rank lookup, summary scores as float sums in query
// order, as score_summaries_typed<uint16_t> adds them), over codes held as
// exact floats; a row's score is the int64 sum of its code products turned to
// float once (dot_product_uint16_dense, utils/distance.h:49-57).
template <bool SQ, bool DENSE, int BLOCK, bool W16 = false>
__device__ __forceinline__ void search_query(const SearchArgs& a, const LdsPlan& P, int64_t q, unsigned char* smem) {
    static_assert(SQ || !DENSE, "the dense code table holds uint8 codes");
    static_assert(!(SQ && W16), "one code width");
    static_assert(BLOCK % 64 == 0 && BLOCK <= 1024, "whole wavefronts, at most 16");
    constexpr int kBlock = BLOCK;
    constexpr int kWaves = BLOCK / 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    uint32_t* qbits = reinterpret_cast<uint32_t*>(smem + P.o_qbits);
    uint16_t* qpref = reinterpret_cast<uint16_t*>(smem + P.o_qpref);
    uint16_t* eterm = reinterpret_cast<uint16_t*>(smem + P.o_eterm);
    uint16_t* erank = reinterpret_cast<uint16_t*>(smem + P.o_erank);
    float* evalf = reinterpret_cast<float*>(smem + P.o_eval);
    int32_t* rlast = reinterpret_cast<int32_t*>(smem + P.o_rlast);
    float* rval = reinterpret_cast<float*>(smem + P.o_rval);
    uint32_t* rsum = reinterpret_cast<uint32_t*>(smem + P.o_rsum);
    int32_t* rcol = reinterpret_cast<int32_t*>(smem + P.o_rcol);
    int32_t* flist = reinterpret_cast<int32_t*>(smem + P.o_flist);
    uint16_t* cuts = reinterpret_cast<uint16_t*>(smem + P.o_cuts);
    int32_t* cstart = reinterpret_cast<int32_t*>(smem + P.o_cstart);
    float* score = reinterpret_cast<float*>(smem + P.o_score);
    uint32_t* sacc = reinterpret_cast<uint32_t*>(smem + P.o_score);
    uint16_t* order = reinterpret_cast<uint16_t*>(smem + P.o_order);
    unsigned char* uni = smem + P.o_uni;
    float* tks = reinterpret_cast<float*>(smem + P.o_tks);
    int32_t* tki = reinterpret_cast<int32_t*>(smem + P.o_tki);
    Seg* seg = reinterpret_cast<Seg*>(smem + P.o_seg);
    int32_t* misc = reinterpret_cast<int32_t*>(smem + P.o_misc);
    int32_t* wsum = reinterpret_cast<int32_t*>(smem + P.o_wsum);
    int32_t* cmeta = reinterpret_cast<int32_t*>(smem + P.o_chunk);
    int64_t* qstat = reinterpret_cast<int64_t*>(smem + P.o_qstat);  // kStatFields
    uint8_t* qcode = smem;  // DENSE only: offset 0
    // the two chunk staging buffers live in the union during the chunk loop;
    // chunk i uses staging buffer i & 1 and piece list / meta slot i % 3
    constexpr int kBud = chunk_budget(DENSE, BLOCK);
    auto chunk_at = [&](int b, int m) {
        int32_t* base = reinterpret_cast<int32_t*>(uni) + b * (5 * kBud + 128);
        Chunk c;
        c.id = base;
        c.len = base + kBud;
        c.rs = base + 2 * kBud;
        c.isc = reinterpret_cast<uint32_t*>(base + 3 * kBud);
        c.sc = reinterpret_cast<float*>(base + 3 * kBud);
        c.wbase = base + 4 * kBud;
        c.seg = seg + m * kSegMax;
        c.meta = cmeta + m * kChunkMetaCount;
        return c;
    };

    const DevIndexView& ix = a.ix;
    const int64_t s0 = a.q_indptr[q];
    const int qn = static_cast<int>(a.q_indptr[q + 1] - s0);
    const int k = a.ks != nullptr ? a.ks[q] : a.k_max;
    float* out_d = a.out_d + q * a.k_max;
    int32_t* out_l = a.out_l + q * a.k_max;
    const uint32_t* sel = a.sel_all;
    bool bad_sel = false;
    if (a.sel_of_query != nullptr) {
        const int s = a.sel_of_query[q];
        bad_sel = s >= a.n_bitmaps;
        sel = s >= 0 && !bad_sel ? a.sel_bitmaps[s] : nullptr;
    }
    const float hf = a.heap_factor;

    if (bad_sel || qn > kQMax || k <= 0 || k > P.kmax) {
        if (tid == 0)
            atomicCAS(a.error_flag, 0, bad_sel ? kDevBadSelector : qn > kQMax ? kDevQueryTooLong : kDevBadK);
        for (int i = tid; i < a.k_max; i += kBlock) {
            out_d[i] = -1.0F;
            out_l[i] = -1;
        }
        return;
    }

    // ---- query setup ----
    if (tid < kMiscCount) misc[tid] = 0;
    if (tid < kStatFields) qstat[tid] = 0;
    const QueryTables QT{qbits, qpref, eterm, erank, evalf, rlast, rval, rsum, qcode, wsum, &misc[kMiscErr],
                         &misc[kMiscNu]};
    if (!load_query_lds<SQ, DENSE, BLOCK, W16>(a, s0, qn, P.nw, QT)) {
        if (tid == 0) atomicCAS(a.error_flag, 0, kDevTermOutOfRange);
        for (int i = tid; i < a.k_max; i += kBlock) {
            out_d[i] = -1.0F;
            out_l[i] = -1;
        }
        return;
    }
    const QueryLds QL{qbits, qpref, rval};
    const int nu = misc[kMiscNu];
    if (tid == 0) {
        // top_k_tokens (utils/vector_process.h:18-30) on codes (SQ) or raw values.
        KeyItem* heap = reinterpret_cast<KeyItem*>(uni);
        misc[kMiscNcut] = topk_tokens(
            qn, a.cut, [&](int i) { return evalf[i]; }, [&](int i) { return static_cast<uint32_t>(eterm[i]); },
            heap, cuts);
    }
    __syncthreads();

    // Traversal counters and phase ticks (stats runs only) are kept in LDS by
    // thread 0, which runs every counted step (wave 0 plans and replays).
    const bool timing = a.stats != nullptr;
    auto stat_add = [&](int f, int64_t v) {
        if (timing && tid == 0) qstat[f] += v;
    };
    uint64_t t_last = timing ? wall_clock64() : 0;
    auto tick = [&](int slot) {
        if (timing && tid == 0) {
            const uint64_t now = wall_clock64();
            qstat[kStatTicksSetup + slot] += static_cast<int64_t>(now - t_last);
            t_last = now;
        }
    };
    tick(0);
    const int ncut = misc[kMiscNcut];
    bool first_list = true;

    for (int ci = 0; ci < ncut; ++ci) {
        const uint32_t term = cuts[ci];
        if (term >= static_cast<uint32_t>(ix.n_lists)) continue;
        stat_add(kStatLists, 1);
        const int32_t lb = ix.list_cluster_base[term];
        const int nc = ix.list_cluster_base[term + 1] - lb;
        if (nc == 0) {
            first_list = false;
            continue;
        }
        const int64_t dbase = ix.cluster_doc_ptr[lb];
        // block-contiguous row copies: word offsets are kept relative to the
        // list's first row
        const bool blk = SQ && ix.row_ref != nullptr;
        const int64_t lwb = blk ? blk_ref_first(ix.row_ref[dbase]) : 0;
        for (int j = tid; j <= nc; j += kBlock) cstart[j] = static_cast<int32_t>(ix.cluster_doc_ptr[lb + j] - dbase);
        const int32_t tb = ix.list_term_base[term];
        const int32_t te = ix.list_term_base[term + 1];
        if (te - tb > 1) stat_add(kStatProbeBytes, static_cast<int64_t>(qn) * 2 * (32 - __clz(te - tb - 1)));
        for (int r = tid; r < nu; r += kBlock) rcol[r] = -1;
        if constexpr (SQ) {
            for (int c = tid; c < nc; c += kBlock) sacc[c] = 0;
        } else {
            for (int c = tid; c < nc; c += kBlock) score[c] = 0.0F;
        }
        if (tid == 0) {
            misc[kMiscNf] = 0;
            misc[kMiscStatFound] = 0;
            misc[kMiscStatCsc] = 0;
        }
        __syncthreads();
        // ---- summary columns matching the query ----
        match_summary_columns<BLOCK>(ix, term, tb, te, qn, eterm, erank, QL, rcol);
        __syncthreads();
        if (a.stats != nullptr) {
            // query entries found in this list and their summary entries (a
            // repeated term counts once per entry, as the reference probes it)
            int fnd = 0, csc = 0;
            for (int i = tid; i < qn; i += kBlock) {
                const int col = rcol[erank[i]];
                if (col >= 0) {
                    ++fnd;
                    csc += static_cast<int>(ix.sum_term_ptr[col + 1] - ix.sum_term_ptr[col]);
                }
            }
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) {
                fnd += __shfl_xor(fnd, d, 64);
                csc += __shfl_xor(csc, d, 64);
            }
            if (lane == 0 && fnd != 0) {
                atomicAdd(&misc[kMiscStatFound], fnd);
                atomicAdd(&misc[kMiscStatCsc], csc);
            }
        }
        if constexpr (SQ) {
            for (int r = tid; r < nu; r += kBlock)
                if (rcol[r] >= 0) flist[atomicAdd(&misc[kMiscNf], 1)] = r;
        } else {
            // found query entries in CSR order (the order the reference adds them in)
            if (wave == 0) {
                int nf = 0;
                for (int b = 0; b < qn; b += 64) {
                    const int i = b + lane;
                    const bool f = i < qn && rcol[erank[i]] >= 0;
                    const uint64_t m = ballot(f);
                    if (f) flist[nf + __popcll(m & ((1ULL << lane) - 1))] = i;
                    nf += __popcll(m);
                }
                if (lane == 0) misc[kMiscNf] = nf;
            }
        }
        __syncthreads();
        const int nf = misc[kMiscNf];
        stat_add(kStatTermsFound, misc[kMiscStatFound]);
        stat_add(kStatCscEntries, misc[kMiscStatCsc]);
        if constexpr (SQ) {
            accumulate_summaries_sq<BLOCK>(ix, tb, te, nf, flist, rcol, rsum, uni, sacc);
            for (int c = tid; c < nc; c += kBlock) score[c] = static_cast<float>(sacc[c]);
        } else {
            const float* sv = static_cast<const float*>(ix.sum_value);
            const uint16_t* sv16 = static_cast<const uint16_t*>(ix.sum_value);
            float* tile = reinterpret_cast<float*>(uni);
            const int per = max(1, (P.uni_bytes / 4) / nc);
            for (int f0 = 0; f0 < nf; f0 += per) {
                const int f1 = min(nf, f0 + per);
                for (int x = tid; x < (f1 - f0) * nc; x += kBlock) tile[x] = 0.0F;
                __syncthreads();
                for (int f = f0 + wave; f < f1; f += kWaves) {
                    const int i = flist[f];
                    const int col = rcol[erank[i]];
                    const float qv = evalf[i];
                    const int64_t e0 = ix.sum_term_ptr[col], e1 = ix.sum_term_ptr[col + 1];
                    float* row = tile + (f - f0) * nc;
                    for (int64_t e = e0 + lane; e < e1; e += 64)
                        row[ix.sum_cluster[e]] = qv * (W16 ? static_cast<float>(sv16[e]) : sv[e]);
                }
                __syncthreads();
                for (int c = tid; c < nc; c += kBlock) {
                    float s = score[c];
                    for (int f = 0; f < f1 - f0; ++f) s += tile[f * nc + c];
                    score[c] = s;
                }
                __syncthreads();
            }
        }
        __syncthreads();
        tick(1);
        // ---- visiting order (seismic_common.h:105-115) ----
        if (first_list) {
            KeyItem* A = reinterpret_cast<KeyItem*>(uni);
            for (int c = tid; c < nc; c += kBlock) A[c] = KeyItem{score[c], static_cast<uint32_t>(c)};
            __syncthreads();
            uint32_t* leaf_bits = reinterpret_cast<uint32_t*>(flist);  // free until the next list
            // every wave takes ranges of each recursion level (one wave alone
            // took 84 us per query at 8.8M docs while the others waited)
            uint16_t* Lpos = reinterpret_cast<uint16_t*>(A + P.mc);
            introsort_desc_block<BLOCK>(A, nc, Lpos, Lpos + P.mc, leaf_bits, reinterpret_cast<int4*>(uni + sort_rng_offset(P.mc)),
                                        sort_range_cap(P.mc), &misc[kMiscSortCnt]);
            __syncthreads();
            introsort_final_pass(A, nc, leaf_bits, order, tid, kBlock);
        } else {
            for (int c = tid; c < nc; c += kBlock) order[c] = static_cast<uint16_t>(c);
        }
        if (tid == 0) {
            misc[kMiscPos] = 0;
            misc[kMiscPoff] = 0;
            misc[kMiscEnd] = nc;
            misc[kMiscDone] = 0;
            misc[kMiscBreakPending] = 0;
        }
        __syncthreads();
        tick(2);

        // ---- speculative chunks + exact replay, software-pipelined ----
        // Chunk i is staged in buffer i & 1. While every wave scores chunk i
        // (windows handed out by an LDS counter), wave 0 replays chunk i - 1
        // exactly, plans chunk i + 1 against the k-th score that replay leaves,
        // gathers and scans it, and then joins the scoring. A plan made against
        // an older (lower) k-th score only takes more clusters, never fewer:
        // the replay's exact test decides every visit, so results and counters
        // are those of the sequential traversal.
        auto plan_chunk = [&](const Chunk& C, int tkn_now) {
            const bool full = tkn_now == k;
            const float kth = full ? tks[k - 1] : 0.0F;
            int pos = misc[kMiscPos], poff = misc[kMiscPoff], end = misc[kMiscEnd];
            int nseg = 0, total = 0;
            bool brk = false;
            // passes that take no piece (every cluster pruned) go on until one
            // does or the list is exhausted; poff > 0 implies pos < end
            while (pos < end) {
                const int p = pos + lane;
                bool valid = p < end;
                const int c = valid ? order[p] : 0;
                const int ctot = valid ? cstart[c + 1] - cstart[c] : 0;
                const bool first_piece = !(lane == 0 && poff > 0);
                bool fail = valid && first_piece && full && (score[c] * hf < kth);
                int new_end = end;
                if (first_list) {
                    const uint64_t fm = ballot(fail);
                    if (fm != 0) {
                        const int fl = __ffsll(static_cast<unsigned long long>(fm)) - 1;
                        valid = valid && lane < fl;
                        fail = false;
                        new_end = pos + fl;
                    }
                }
                const bool pass = valid && !fail;
                const int cnt = pass ? ctot - (lane == 0 ? poff : 0) : 0;
                const int incl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(cnt)));
                const int excl = incl - cnt;
                const bool stopper = pass && (excl + cnt > kBud);
                const uint64_t sm = ballot(stopper);
                const int stl = sm ? __ffsll(static_cast<unsigned long long>(sm)) - 1 : 64;
                const bool partial = lane == stl && excl < kBud;
                const bool is_seg = pass && (lane < stl || partial);
                const uint64_t segm = ballot(is_seg);
                if (is_seg) {
                    Seg s;
                    s.doc_begin = cstart[c] + (lane == 0 ? poff : 0);
                    s.cluster = static_cast<uint16_t>(c);
                    s.pos = static_cast<uint16_t>(p);
                    s.count = static_cast<uint16_t>(partial ? kBud - excl : cnt);
                    s.slot = static_cast<uint16_t>(excl);
                    s.test = first_piece ? 1 : 0;
                    s.pad = 0;
                    C.seg[__popcll(segm & ((1ULL << lane) - 1))] = s;
                }
                // consumed non-first-list failures were prune tests too
                stat_add(kStatClustersTested, __popcll(ballot(valid && fail && lane < stl)));
                int new_pos, new_poff;
                if (sm) {
                    new_pos = pos + stl;
                    const int before = (stl == 0 ? poff : 0);
                    new_poff = excl < kBud ? before + (kBud - excl) : before;
                    // `excl`/`poff` above are lane stl's; broadcast them
                    new_poff = __shfl(new_poff, stl, 64);
                } else {
                    new_pos = min(pos + 64, new_end);
                    new_poff = 0;
                }
                brk = brk || new_end != end;
                pos = new_pos;
                poff = new_poff;
                end = new_end;
                nseg = __popcll(segm);
                total = min(__builtin_amdgcn_readlane(incl, 63), kBud);
                if (nseg > 0) break;
            }
            if (lane == 0) {
                misc[kMiscPos] = pos;
                misc[kMiscPoff] = poff;
                misc[kMiscEnd] = end;
                if (brk) misc[kMiscBreakPending] = 1;
                C.meta[kChunkNseg] = nseg;
                C.meta[kChunkTotal] = nseg > 0 ? total : 0;
            }
            __builtin_amdgcn_wave_barrier();
        };

        // Doc ids, selector and row ranges of the chunk's slots, by threads
        // t0, t0 + nt, ...; kUnroll slots per thread are resolved before their
        // loads are issued. Out-of-range slots load the last slot's doc again.
        // `overlap` runs exactly once, while the first round's doc-id and row
        // loads are in flight (rounds are uniform over the threads); it must not
        // touch the slots being gathered.
        // `abandoned` (wave 0 only) is set by `overlap` when the other waves
        // may still be reading the staging buffer the slots go to; the slots
        // are then left untouched.
        bool abandoned = false;
        auto gather_slots = [&](auto unroll_tag, const Chunk& C, int t0, int nt, auto&& overlap) {
            constexpr int kUnroll = decltype(unroll_tag)::value;
            const int nseg = C.meta[kChunkNseg], total = C.meta[kChunkTotal];
            if (total == 0) overlap();
            for (int r0 = 0; r0 < total && !abandoned; r0 += nt * kUnroll) {
                const int j0 = r0 + t0;
                int64_t pp[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const int j = min(j0 + u * nt, total - 1);
                    int lo = 0, hi = nseg - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (C.seg[mid].slot <= j)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    pp[u] = dbase + C.seg[lo].doc_begin + (j - C.seg[lo].slot);
                }
                int32_t d[kUnroll];
                int64_t ref[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    d[u] = ix.docs[pp[u]];
                    ref[u] = blk ? ix.row_ref[pp[u]] : 0;
                }
                if (r0 == 0) {
                    overlap();
                    if (abandoned) break;  // wave-uniform
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const int j = j0 + u * nt;
                    const bool keep = sel == nullptr || ((sel[d[u] >> 5] >> (d[u] & 31)) & 1u);
                    int32_t rs, len;
                    if (blk) {
                        // rs in 8ths of words, as for forward-index rows; len
                        // packs the word count over the nnz
                        rs = static_cast<int32_t>((blk_ref_first(ref[u]) - lwb) << 3);
                        len = (blk_ref_words(ref[u]) << 16) | blk_ref_nnz(ref[u]);
                    } else {
                        rs = ix.indptr[d[u]];
                        len = ix.indptr[d[u] + 1] - rs;
                    }
                    if (j < total) {
                        C.id[j] = d[u];
                        C.rs[j] = rs;
                        C.len[j] = keep ? len : -1;
                    }
                }
            }
        };

        // Word starts of the staged rows (wave 0). Rows are read as aligned
        // 8-element words; wbase[j] is doc j's first word within the chunk,
        // wbase[total] the chunk's word count, followed by sentinels above every
        // word index (the boundary search reads up to wbase[total + 64]). Block
        // rows keep, in rs, the doc's first word minus its first chunk word.
        // Also clears the integer sums and the window counter.
        auto scan_words = [&](const Chunk& C) {
            const int total = C.meta[kChunkTotal];
            int run = 0;
            if constexpr (SQ) {
                for (int j0 = 0; j0 < total; j0 += 64) {
                    const int j = j0 + lane;
                    int nwj = 0;
                    if (j < total) {
                        const int len = C.len[j];
                        // block rows: 32-byte pairs of words
                        if (len > 0) nwj = blk ? (len >> 17) : ((C.rs[j] + len + 7) >> 3) - (C.rs[j] >> 3);
                        C.isc[j] = 0;
                    }
                    const int incl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(nwj)));
                    if (j < total) {
                        C.wbase[j] = run + incl - nwj;
                        if (blk) C.rs[j] = (C.rs[j] >> 4) - (run + incl - nwj);
                    }
                    run += __builtin_amdgcn_readlane(incl, 63);
                }
                C.wbase[total + lane] = lane == 0 ? run : 0x7FFFFFFF;
                C.wbase[total + 64 + lane] = 0x7FFFFFFF;
            }
            if (lane == 0) {
                C.meta[kChunkWords] = run;
                C.meta[kChunkNext] = 0;
            }
            __builtin_amdgcn_wave_barrier();
        };

        // ---- score the staged rows (every wave) ----
        auto score_chunk = [&](const Chunk& C) {
            const int total = C.meta[kChunkTotal];
            if constexpr (SQ) {
                // The chunk's words, doc after doc, form windows of 64 words, one
                // word per lane, so every load is coalesced; a wave takes
                // kWinPerStep consecutive windows per step from the counter and
                // keeps their loads in flight together. A window's first doc is
                // found with two ballots over wbase (every kBud/64-th start, then
                // the kBud/64 starts after the last one <= the window's first
                // word); its lanes then find their doc from the doc boundaries
                // inside the window (a wave-uniform loop). A DPP prefix sum S
                // over the window yields each doc run's sum: the last lane of a
                // run adds S to its doc and subtracts S from the next run's doc
                // (LDS integer atomics: exact in any order, modulo 2^32, and the
                // true sums fit). For block-contiguous rows a "word" here is a
                // 32-byte pair of 16-byte block words (rows are whole pairs), so
                // the doc search, prefix sum and atomics of a window cover 128
                // block words: with 8-slot block words the per-window work, not
                // HBM, bounded the step. Block rows need no masks (unused slots
                // are zero); lanes without a word read the zero pair past the
                // end. Forward-index rows mask the bytes outside the row.
                constexpr int kWinPerStep = win_per_step<DENSE>();
                constexpr int kS1 = kBud / 64;
                const int W = C.meta[kChunkWords];
                const int nwin = (W + 63) >> 6;
                const int32_t* wbase = C.wbase;
                auto score_windows = [&](auto blk_tag) {
                    constexpr bool BLK = decltype(blk_tag)::value;
                    // block rows: a lane scores one 32-byte pair of words
                    constexpr int kWin = BLK ? kBlkWinPerStep : kWinPerStep;
                    const uint4* idx8 = BLK ? ix.blk_rows + lwb : reinterpret_cast<const uint4*>(ix.indices);
                    const uint2* val8 = reinterpret_cast<const uint2*>(ix.values);  // forward-index rows
                    const int64_t zero_word = BLK ? (ix.blk_words - lwb) >> 1 : 0;
                    // every kS1-th word start, read once per chunk: the first
                    // ballot of each window's doc search needs no LDS read
                    const int wb1 = lane * kS1 < total ? wbase[lane * kS1] : INT32_MAX;
                    // One step: kWin consecutive windows, their lanes' docs,
                    // word addresses and loaded words.
                    struct Step {
                        int jj[kWin], j1s[kWin], lo[kWin], hi[kWin];
                        int64_t ww[kWin];
                        uint4 ii[kWin], ii2[kWin];
                        uint2 vv[kWin];
                    };
                    auto fetch = [&]() {
                        int xs = 0;
                        if (lane == 0) xs = atomicAdd(&C.meta[kChunkNext], kWin);
                        return __shfl(xs, 0, 64);
                    };
                    // 1. each lane's doc and word address. The windows of a
                    // step are independent, so each phase issues its LDS
                    // reads for all of them before any result is used; a
                    // window past the last one repeats the last window and
                    // is masked out (no word, no sums).
                    auto locate = [&](int xs, Step& S) {
                        int sbs[kWin], bns[kWin];
                        // 1a. the starts after every kS1-th one <= the window's first word
#pragma unroll
                        for (int u = 0; u < kWin; ++u) {
                            const int F = min(xs + u, nwin - 1) << 6;
                            sbs[u] = (__popcll(ballot(wb1 <= F)) - 1) * kS1;
                            bns[u] = wbase[sbs[u] + lane];  // sentinels cover sb + 63
                        }
                        // 1b. the window's first doc, then the doc boundaries inside it
#pragma unroll
                        for (int u = 0; u < kWin; ++u) {
                            const bool valid = xs + u < nwin;  // wave-uniform
                            const int F = min(xs + u, nwin - 1) << 6, f = F + lane;
                            const int sb = sbs[u], bn0 = bns[u];
                            // the last j < total with wbase[j] <= F, among sb .. sb + kS1 - 1
                            const uint64_t m2 = ballot(lane < kS1 && sb + lane < total && bn0 <= F);
                            int j = sb + __popcll(m2) - 1;
                            // boundaries after j: lanes j - sb + 1 .. 63 of bn0 hold
                            // wbase[j + 1 .. sb + 63]; a window with more starts
                            // than that reads on
                            const int r0 = j - sb + 1;
                            const uint64_t inside = ballot(lane >= r0 && bn0 <= F + 63);
                            const int nb0 = __popcll(inside);
                            for (int i = 0; i < nb0; ++i) j += f >= __builtin_amdgcn_readlane(bn0, r0 + i);
                            if (r0 + nb0 == 64) {
                                for (int base = sb + 64;; base += 64) {
                                    const int bn = wbase[base + lane];
                                    const uint64_t bm = ballot(bn <= F + 63);
                                    const int nb = __popcll(bm);
                                    for (int i = 0; i < nb; ++i) j += f >= __builtin_amdgcn_readlane(bn, i);
                                    if (nb < 64) break;
                                }
                            }
                            // the doc of word f + 1, from the next lane (a DPP
                            // wavefront shift left by one lane; lane 63 reads 0)
                            const int j1 = __builtin_amdgcn_update_dpp(0, j, 0x130, 0xF, 0xF, false);
                            const bool has = valid && f < W;
                            S.jj[u] = has ? j : -1;
                            // run end (the next word is another doc's, or the window
                            // ends): j1s = the doc whose run follows in this window,
                            // INT32_MAX when none does, -1 when not a run end
                            const bool run_end = j1 != j || lane == 63;
                            S.j1s[u] = !has || !run_end ? -1 : (j1 != j && f + 1 < W && lane != 63) ? j1 : INT32_MAX;
                        }
                        // 1c. row references of the lanes' docs
#pragma unroll
                        for (int u = 0; u < kWin; ++u) {
                            const int F = min(xs + u, nwin - 1) << 6, f = F + lane;
                            const int j = S.jj[u];
                            S.ww[u] = zero_word;
                            S.lo[u] = 0;
                            S.hi[u] = 8;
                            if constexpr (BLK) {
                                const int r = C.rs[max(j, 0)];
                                if (j >= 0) S.ww[u] = f + r;
                            } else {
                                const int jc = max(j, 0);
                                const int rs = C.rs[jc], re = rs + C.len[jc], wb = wbase[jc];
                                if (j >= 0) {
                                    const int w = (rs >> 3) + (f - wb);
                                    S.ww[u] = w;
                                    S.lo[u] = max(rs - (w << 3), 0);
                                    S.hi[u] = min(re - (w << 3), 8);
                                }
                            }
                        }
                    };
                    // 2. all loads of the step back to back, without branches
                    auto load = [&](Step& S) {
#pragma unroll
                        for (int u = 0; u < kWin; ++u) {
                            if constexpr (BLK) {
                                S.ii[u] = idx8[2 * S.ww[u]];
                                S.ii2[u] = idx8[2 * S.ww[u] + 1];
                            } else {
                                S.ii[u] = idx8[S.ww[u]];
                                S.vv[u] = val8[S.ww[u]];
                            }
                        }
                    };
                    // 3. scores (a masked window scores the zero word), then
                    // the windows' prefix sums, then their atomics: three
                    // passes, so the windows' LDS reads and DPP chains
                    // interleave
                    auto finish = [&](Step& S) {
                        uint32_t accs[kWin];
#pragma unroll
                        for (int u = 0; u < kWin; ++u) {
                            const uint32_t tw[4] = {S.ii[u].x, S.ii[u].y, S.ii[u].z, S.ii[u].w};
                            uint32_t acc = 0;
                            auto add = [&](uint32_t term, uint32_t c) {
                                if constexpr (DENSE) {
                                    acc += c * static_cast<uint32_t>(qcode[term]);
                                } else {
                                    bool pr;
                                    const int r = QL.rank_of(term, pr);
                                    const uint32_t qv = static_cast<uint32_t>(rval[r]);
                                    acc += pr ? c * qv : 0U;
                                }
                            };
                            if constexpr (BLK) {
                                if constexpr (DENSE) {
                                    acc = blk_word_dot(S.ii2[u], qcode, blk_word_dot(S.ii[u], qcode, 0U));
                                } else {
                                    // 8 slots of (offset << 4 | code) under each word's base
#pragma unroll
                                    for (int h = 0; h < 2; ++h) {
                                        const uint4 w = h == 0 ? S.ii[u] : S.ii2[u];
                                        const uint32_t tv[4] = {w.x, w.y, w.z, w.w};
                                        const uint32_t base = blk_word_base(tv[0]);
#pragma unroll
                                        for (int t = 0; t < 8; ++t) {
                                            const uint32_t slot = (t & 1) ? (tv[t >> 1] >> 16) : (tv[t >> 1] & 0xFFFFU);
                                            add((t == 0 ? 0U : base) + (slot >> 4), slot & 15U);
                                        }
                                    }
                                }
                            } else {
                                // bytes outside the doc's [rs, re) -- and every byte of a
                                // lane without a word -- are masked; their term ids are
                                // still valid ids, so the lookups stay in bounds
                                uint64_t codes = (static_cast<uint64_t>(S.vv[u].y) << 32) | S.vv[u].x;
                                const bool has = S.jj[u] >= 0;
                                const uint64_t mlo = has ? (~0ULL << (8 * S.lo[u])) : 0ULL;
                                const uint64_t mhi = S.hi[u] < 8 ? ((1ULL << (8 * S.hi[u])) - 1) : ~0ULL;
                                codes &= mlo & mhi;
#pragma unroll
                                for (int t = 0; t < 8; ++t) {
                                    const uint32_t term = (t & 1) ? (tw[t >> 1] >> 16) : (tw[t >> 1] & 0xFFFFu);
                                    add(term, static_cast<uint32_t>(codes >> (8 * t)) & 0xFFu);
                                }
                            }
                            accs[u] = acc;
                        }
#pragma unroll
                        for (int u = 0; u < kWin; ++u) accs[u] = wave_incl_scan_dpp(accs[u]);
#pragma unroll
                        for (int u = 0; u < kWin; ++u) {
                            const uint32_t Sum = accs[u];
                            if (S.j1s[u] >= 0 && Sum != 0) {
                                atomicAdd(&C.isc[S.jj[u]], Sum);
                                if (S.j1s[u] != INT32_MAX) atomicSub(&C.isc[S.j1s[u]], Sum);
                            }
                        }
                    };
                    // Keeping the next step's loads in flight while this step is
                    // scored (software pipelining, 2 or 3 windows per step)
                    // measured 5-7% slower at 8.8M docs: the extra registers
                    // spill and the extra loads in flight congest HBM.
                    while (true) {
                        const int xs = fetch();
                        if (xs >= nwin) break;
                        Step S;
                        locate(xs, S);
                        load(S);
                        finish(S);
                    }
                };
                if (blk)
                    score_windows(std::true_type{});
                else
                    score_windows(std::false_type{});
                // this wave's sums are in (release: its LDS atomics come first)
                if (lane == 0)
                    __hip_atomic_fetch_add(&C.meta[kChunkCheckin], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                // FP32 rows: one doc per lane, 64-slot groups from the counter;
                // each row is summed in order (as the reference's loop does)
                const float* vf = static_cast<const float*>(ix.values);
                const int64_t nnz_end = ix.indptr[ix.n_docs];
                while (true) {
                    int g = 0;
                    if (lane == 0) g = atomicAdd(&C.meta[kChunkNext], 1);
                    g = __shfl(g, 0, 64);
                    if (g * 64 >= total) break;
                    const int j = g * 64 + lane;
                    const int len = j < total ? C.len[j] : -1;
                    if (len < 0) continue;
                    const int32_t rs = C.rs[j];
                    // the row as aligned words of 8 entries (16 bytes of term ids,
                    // 32 of float or 16 of 16-bit values), kRowWords per step,
                    // entries still in row order; a word past the arrays' last
                    // entry is read entry by entry
                    constexpr int kRowWords = 1;
                    const uint4* iw = reinterpret_cast<const uint4*>(ix.indices);
                    const int64_t r0 = rs, r1 = static_cast<int64_t>(rs) + len;
                    const int64_t wend = (r1 + 7) >> 3;
                    if constexpr (W16) {
                        const uint16_t* v16 = static_cast<const uint16_t*>(ix.values);
                        const uint4* vw = reinterpret_cast<const uint4*>(ix.values);
                        uint64_t acc = 0;
                        for (int64_t w0 = r0 >> 3; w0 < wend; w0 += kRowWords) {
                            if ((w0 + kRowWords) * 8 <= nnz_end) {
                                uint4 ti[kRowWords], tv[kRowWords];
#pragma unroll
                                for (int u = 0; u < kRowWords; ++u) {
                                    ti[u] = iw[w0 + u];
                                    tv[u] = vw[w0 + u];
                                }
#pragma unroll
                                for (int u = 0; u < kRowWords; ++u) {
                                    const uint32_t tt[4] = {ti[u].x, ti[u].y, ti[u].z, ti[u].w};
                                    const uint32_t vv[4] = {tv[u].x, tv[u].y, tv[u].z, tv[u].w};
#pragma unroll
                                    for (int e = 0; e < 8; ++e) {
                                        const int64_t pp = (w0 + u) * 8 + e;
                                        if (pp >= r0 && pp < r1) {
                                            const uint32_t t = (e & 1) ? (tt[e >> 1] >> 16) : (tt[e >> 1] & 0xFFFFu);
                                            const uint32_t c = (e & 1) ? (vv[e >> 1] >> 16) : (vv[e >> 1] & 0xFFFFu);
                                            acc += static_cast<uint64_t>(c * static_cast<uint32_t>(QL.value(t)));
                                        }
                                    }
                                }
                            } else {
                                for (int64_t pp = w0 * 8 > r0 ? w0 * 8 : r0; pp < r1; ++pp)
                                    acc += static_cast<uint64_t>(static_cast<uint32_t>(v16[pp]) *
                                                                 static_cast<uint32_t>(QL.value(ix.indices[pp])));
                            }
                        }
                        C.sc[j] = static_cast<float>(acc);
                    } else {
                        const float4* vw = reinterpret_cast<const float4*>(ix.values);
                        float acc = 0.0F;
                        for (int64_t w0 = r0 >> 3; w0 < wend; w0 += kRowWords) {
                            if ((w0 + kRowWords) * 8 <= nnz_end) {
                                uint4 ti[kRowWords];
                                float4 va[kRowWords], vb[kRowWords];
#pragma unroll
                                for (int u = 0; u < kRowWords; ++u) {
                                    ti[u] = iw[w0 + u];
                                    va[u] = vw[2 * (w0 + u)];
                                    vb[u] = vw[2 * (w0 + u) + 1];
                                }
#pragma unroll
                                for (int u = 0; u < kRowWords; ++u) {
                                    const uint32_t tt[4] = {ti[u].x, ti[u].y, ti[u].z, ti[u].w};
                                    const float vv[8] = {va[u].x, va[u].y, va[u].z, va[u].w,
                                                         vb[u].x, vb[u].y, vb[u].z, vb[u].w};
#pragma unroll
                                    for (int e = 0; e < 8; ++e) {
                                        const int64_t pp = (w0 + u) * 8 + e;
                                        if (pp >= r0 && pp < r1) {
                                            const uint32_t t = (e & 1) ? (tt[e >> 1] >> 16) : (tt[e >> 1] & 0xFFFFu);
                                            acc += vv[e] * QL.value(t);
                                        }
                                    }
                                }
                            } else {
                                for (int64_t pp = w0 * 8 > r0 ? w0 * 8 : r0; pp < r1; ++pp)
                                    acc += vf[pp] * QL.value(ix.indices[pp]);
                            }
                        }
                        C.sc[j] = acc;
                    }
                }
                if (lane == 0)
                    __hip_atomic_fetch_add(&C.meta[kChunkCheckin], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        };

        // ---- exact replay of a scored chunk (wave 0) ----
        // Segments are replayed in visiting order: the prune test against the
        // exact current k-th, then the segment's docs are merged into the
        // sorted top-k one 64-slot word at a time, lane l holding slot 64w + l.
        // Inserting a batch one by one with the reference's reject rule keeps
        // the k best of (top-k U batch) under (score desc, doc id asc) -- an
        // order-free result -- so the batch is merged in parallel: each
        // candidate's rank among the top-k (ballots against the top-k held in
        // registers, or a binary search when it has more than 128 entries; an
        // equal id is the same doc seen in an earlier list) plus its rank
        // within the batch. A cluster split over chunks is tested at its first
        // piece; when that test fails (on a later list) its other pieces are
        // skipped too (`cluster_failed` carries the outcome across chunks).
        // Returns true at a first-list break.
        bool cluster_failed = false;
        auto replay_chunk = [&](const Chunk& C, int& tkn) -> bool {
            auto merge = [&](bool in, int j) {
                float sc = 0.0F;
                if (in) {
                    if constexpr (SQ)
                        sc = static_cast<float>(C.isc[j]);
                    else
                        sc = C.sc[j];
                }
                merge_topk_wave(in, sc, in ? C.id[j] : 0, k, tkn, tks, tki, lane);
            };
            const int nseg = C.meta[kChunkNseg];
            for (int si = 0; si < nseg; ++si) {
                const Seg s = C.seg[si];
                if (s.test) {
                    stat_add(kStatClustersTested, 1);
                    cluster_failed = tkn == k && score[s.cluster] * hf < tks[k - 1];
                    if (cluster_failed) {
                        if (first_list) return true;
                        continue;
                    }
                    stat_add(kStatClustersOpened, 1);
                } else if (cluster_failed) {
                    continue;
                }
                stat_add(kStatDocsRead, s.count);
                const int send = s.slot + s.count;
                for (int w = s.slot >> 6; w < (send + 63) >> 6; ++w) {
                    const int j = w * 64 + lane;
                    const int len = (j >= s.slot && j < send) ? C.len[j] : -1;
                    if (timing) {
                        stat_add(kStatDocsScored, __popcll(ballot(len >= 0)));
                        const int nnz = len < 0 ? 0 : blk ? (len & 0xFFFF) : len;
                        const uint32_t incl = wave_incl_scan_dpp(static_cast<uint32_t>(nnz));
                        stat_add(kStatNnzScored,
                                 static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63)));
                    }
                    merge(len >= 0, j);
                }
            }
            return false;
        };

        // Spins (uniform over the wave, s_sleep between polls) until *p == v.
        // After a.spin_max polls (about a quarter second by default) it reports
        // kDevChunkWait and returns false; once that is reported, every wait
        // of the launch returns false at once, so a launch always ends.
        auto wait_for = [&](const int32_t* p, int v) {
            for (int it = 0; lds_load_acquire(p) != v; ++it) {
                if ((it & 1023) == 0 &&
                    __hip_atomic_load(a.error_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == kDevChunkWait)
                    return false;
                if (it >= a.spin_max) {
                    if (lane == 0) atomicCAS(a.error_flag, 0, kDevChunkWait);
                    return false;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            return true;
        };

        // prologue: chunk 0 planned by wave 0, gathered by every wave
        {
            const Chunk C0 = chunk_at(0, 0);
            if (wave == 0) {
                plan_chunk(C0, misc[kMiscTkN]);
                if (lane < 3) cmeta[lane * kChunkMetaCount + kChunkStamp] = -1;
                if (lane == 0) C0.meta[kChunkCheckin] = 0;
            }
            __syncthreads();
            tick(3);
            gather_slots(std::integral_constant<int, (kBud + kBlock - 1) / kBlock>{}, C0, tid, kBlock, [] {});
            __syncthreads();
            if (wave == 0) {
                scan_words(C0);
                if (lane == 0) {
                    C0.meta[kChunkEnd] = C0.meta[kChunkNseg] == 0 ? 1 : 0;
                    C0.meta[kChunkStamp] = 0;
                }
            }
            __syncthreads();
            tick(4);
        }
        // Chunk i sits in staging buffer i & 1 and slot i % 3, published by its
        // stamp. Every wave scores chunk i as soon as it is published and then
        // checks in on it; no barrier separates the chunks. Wave 0, in its
        // iteration i, plans chunk i + 1 (against the k-th score after chunk
        // i - 2's replay) and issues its gather loads; meanwhile it waits for
        // every wave's check-in on chunk i - 1 and replays that chunk; then it
        // stages and scans chunk i + 1, publishes it (or an end marker after a
        // first-list break or the last cluster), and joins the scoring of
        // chunk i. The check-in wait keeps every wave within one chunk of wave
        // 0, so the buffers chunk i + 1 reuses (those of chunks i - 1 and
        // i - 2) are free.
        if (wave == 0) {
            bool have_prev = false, stop = false;
            // wave 0's control work (replay, plan, gather, scan) is the chain
            // every other wave waits on, so wave 0 issues first on its SIMD
            // (+0.9% batched throughput at 8.8M docs; dropping back to the
            // scoring waves' priority while it scores measured 0.1% lower)
            __builtin_amdgcn_s_setprio(2);
            for (int i = 0, cb = 0, cm = 0;; ++i, cb ^= 1, cm = cm == 2 ? 0 : cm + 1) {
                const Chunk cur = chunk_at(cb, cm);
                const Chunk prev = chunk_at(cb ^ 1, cm == 0 ? 2 : cm - 1);
                const Chunk next = chunk_at(cb ^ 1, cm == 2 ? 0 : cm + 1);
                int tkn = misc[kMiscTkN];
                auto replay_prev = [&] {
                    if (!have_prev) return;
                    tick(4);
                    // A wait that gave up (error 5 is set, the launch's
                    // results are void) leaves waves that may still score
                    // chunk i - 1 from the staging buffer chunk i + 1 would
                    // reuse: stop without gathering or scanning into it and
                    // publish an end marker instead.
                    if (!wait_for(&prev.meta[kChunkCheckin], kWaves)) {
                        stop = true;
                        abandoned = true;
                    }
                    if (!stop) {
                        stop = replay_chunk(prev, tkn);
                        if (lane == 0) misc[kMiscTkN] = tkn;
                        __builtin_amdgcn_wave_barrier();
                    }
                    tick(6);
                };
                if (cur.meta[kChunkEnd] != 0) {
                    replay_prev();
                    // every planned chunk is replayed; a break found by the
                    // planner is the prune test at the position where the visit stops
                    if (!stop && misc[kMiscBreakPending]) stat_add(kStatClustersTested, 1);
                    break;
                }
                plan_chunk(next, tkn);
                if (lane == 0) next.meta[kChunkCheckin] = 0;
                tick(3);
                gather_slots(std::integral_constant<int, (kBud / 64 < 8 ? kBud / 64 : 8)>{}, next, lane, 64,
                             replay_prev);
                if (!abandoned) scan_words(next);
                if (lane == 0) {
                    next.meta[kChunkEnd] = (stop || next.meta[kChunkNseg] == 0) ? 1 : 0;
                    lds_store_release(&next.meta[kChunkStamp], i + 1);
                }
                __builtin_amdgcn_wave_barrier();
                tick(4);
                score_chunk(cur);
                tick(5);
                have_prev = true;
            }
            __builtin_amdgcn_s_setprio(0);
        } else {
            for (int i = 0, cb = 0, cm = 0;; ++i, cb ^= 1, cm = cm == 2 ? 0 : cm + 1) {
                const Chunk cur = chunk_at(cb, cm);
                if (!wait_for(&cur.meta[kChunkStamp], i)) break;
                if (cur.meta[kChunkEnd] != 0) break;
                score_chunk(cur);
            }
        }
        __syncthreads();
        first_list = false;
        __syncthreads();
    }

    // ---- results: best first, (-1.0f, -1) padding ----
    const int tkn = misc[kMiscTkN];
    for (int i = tid; i < a.k_max; i += kBlock) {
        float d = -1.0F;
        int32_t l = -1;
        if (i < k && i < tkn) {
            d = (SQ || W16) ? tks[i] * a.decode_scale : tks[i];
            l = tki[i];
        }
        out_d[i] = d;
        out_l[i] = l;
    }
    if (timing && tid == 0) {
        int64_t* s = a.stats + q * kStatFields;
        for (int f = 0; f < kStatFields; ++f) s[f] = qstat[f];
    }
}

// One workgroup per query: query query_base + blockIdx.x. Both kernels are held
// to 128 VGPRs (4 waves per SIMD: two 512-lane or four 256-lane queries per CU).
template <bool SQ, bool DENSE, int BLOCK, bool W16 = false>
__global__ __launch_bounds__(BLOCK, 4) void seismic_search_kernel(SearchArgs a, LdsPlan P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t q = a.query_base + blockIdx.x;
    if (q >= a.n_queries) return;
    search_query<SQ, DENSE, BLOCK, W16>(a, P, q, smem);
}

// Persistent workgroups (as many as fit on the device at once): each takes the
// next query index from the device counter a.queue (zero at launch) until all
// n_queries are handed out; a.order, when set, maps the index to a query (for
// longest-first scheduling). Every workgroup exits once the counter passes n.
template <bool SQ, bool DENSE, int BLOCK, bool W16 = false>
__global__ __launch_bounds__(BLOCK, 4) void seismic_search_persistent_kernel(SearchArgs a, LdsPlan P) {
    extern __shared__ __align__(16) unsigned char smem[];
    int32_t* slot = reinterpret_cast<int32_t*>(smem + P.o_queue);
    while (true) {
        if (threadIdx.x == 0) *slot = atomicAdd(a.queue, 1);
        __syncthreads();
        const int64_t i = *slot;
        __syncthreads();
        if (i >= a.n_queries) break;
        const int64_t q = a.order != nullptr ? a.order[i] : i;
        search_query<SQ, DENSE, BLOCK, W16>(a, P, q, smem);
        __syncthreads();
    }
}

namespace {

LdsPlan make_plan(int dim, int max_clusters, int kmax, bool sq, bool dense, int block) {
    LdsPlan P{};
    P.nw = (dim + 31) / 32;
    P.mc = max(max_clusters, 1);
    P.kmax = kmax;
    P.budget = chunk_budget(dense, block);
    const int staging = 2 * (5 * P.budget + 128) * 4;  // two chunk buffers
    // KeyItem[mc] + two uint16[mc] position lists + two levels of sort ranges
    const int sortbuf = sort_rng_offset(P.mc) + 2 * sort_range_cap(P.mc) * 16;
    const int tile = sq ? 0 : kTileFloats * 4;
    P.uni_bytes = max(max(staging, sortbuf), max(tile, kQMax * 8));
    int o = 0;
    auto take = [&](int bytes) {
        const int at = o;
        o += (bytes + 15) & ~15;
        return at;
    };
    if (dense) take(dim);  // the dense code table must sit at offset 0
    P.o_qbits = take(P.nw * 4);
    P.o_qpref = take(P.nw * 2);
    P.o_eterm = take(kQMax * 2);
    P.o_erank = take(kQMax * 2);
    P.o_eval = take(kQMax * 4);
    P.o_rlast = take(kQMax * 4);  // query setup only; shared with rcol (per list)
    P.o_rcol = P.o_rlast;
    P.o_rval = take(kQMax * 4);
    P.o_rsum = take(kQMax * 4);
    P.o_flist = take(kQMax * 4);
    P.o_cuts = take(kQMax * 2);
    P.o_cstart = take((P.mc + 1) * 4);
    P.o_score = take(P.mc * 4);
    P.o_order = take(P.mc * 2);
    P.o_uni = take(P.uni_bytes);
    P.o_tks = take(kmax * 4);
    P.o_tki = take(kmax * 4);
    P.o_seg = take(3 * kSegMax * static_cast<int>(sizeof(Seg)));
    P.o_misc = take(kMiscCount * 4);
    P.o_wsum = take(kMaxWaves * 4);
    P.o_queue = take(16);
    P.o_chunk = take(3 * kChunkMetaCount * 4);
    P.o_qstat = take(kStatFields * 8);
    P.total = o;
    return P;
}

}  // namespace

// ExactMatcher::single_query (exact_matcher.h:23-41): every id of a small
// enumerable selector is scored against the dense query and the <= k results
// are written best first, ties by doc id. One workgroup per query.
template <bool SQ, bool W16 = false>
__global__ __launch_bounds__(kBlock) void exact_match_kernel(ExactArgs a, int nw) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* qbits = reinterpret_cast<uint32_t*>(smem);
    uint16_t* qpref = reinterpret_cast<uint16_t*>(qbits + nw);
    float* rval = reinterpret_cast<float*>(smem + ((nw * 6 + 15) & ~15));
    int32_t* rlast = reinterpret_cast<int32_t*>(rval + kQMax);
    float* sc = reinterpret_cast<float*>(rlast + kQMax);
    int32_t* wsum = reinterpret_cast<int32_t*>(sc + kQMax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = blockIdx.x;
    if (q >= a.n_queries) return;
    const int64_t s0 = a.q_indptr[q];
    const int qn = static_cast<int>(a.q_indptr[q + 1] - s0);
    const DevIndexView& ix = a.ix;
    for (int w = tid; w < nw; w += kBlock) qbits[w] = 0;
    for (int r = tid; r < qn; r += kBlock) rlast[r] = -1;
    __syncthreads();
    for (int i = tid; i < qn; i += kBlock) {
        const uint32_t t = a.q_indices[s0 + i];
        atomicOr(&qbits[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    {
        const int per = (nw + kBlock - 1) / kBlock;
        const int w0 = tid * per;
        int local = 0;
        for (int w = w0; w < w0 + per && w < nw; ++w) local += __popc(qbits[w]);
        const int incl = wave_incl_scan(local, lane);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int run = incl - local;
        for (int w = 0; w < wave; ++w) run += wsum[w];
        for (int w = w0; w < w0 + per && w < nw; ++w) {
            qpref[w] = static_cast<uint16_t>(run);
            run += __popc(qbits[w]);
        }
    }
    __syncthreads();
    const QueryLds QL{qbits, qpref, rval};
    for (int i = tid; i < qn; i += kBlock) {
        bool p;
        atomicMax(&rlast[QL.rank_of(a.q_indices[s0 + i], p)], i);
    }
    __syncthreads();
    // SparseVectors::get_dense_vector: the last entry of a repeated term wins.
    for (int r = tid; r < qn; r += kBlock) {
        const int i = rlast[r];
        if (i < 0) continue;
        const float v = a.q_values[s0 + i];
        if constexpr (SQ || W16) {
            constexpr float kMax = W16 ? 65535.0F : 255.0F;
            float x = (v - a.q_vmin) * a.q_scale;
            x = (x < kMax) ? x : kMax;
            x = (0.0F < x) ? x : 0.0F;
            rval[r] = static_cast<float>(static_cast<uint32_t>(roundf(x)));
        } else {
            rval[r] = v;
        }
    }
    __syncthreads();
    for (int j = tid; j < a.n_ids; j += kBlock) {
        const int32_t d = a.ids[j];
        const int32_t r0 = ix.indptr[d], r1 = ix.indptr[d + 1];
        if constexpr (SQ) {
            const uint8_t* vv = static_cast<const uint8_t*>(ix.values);
            int acc = 0;
            for (int p = r0; p < r1; ++p) acc += static_cast<int>(vv[p]) * static_cast<int>(QL.value(ix.indices[p]));
            sc[j] = static_cast<float>(acc);
        } else if constexpr (W16) {
            const uint16_t* vv = static_cast<const uint16_t*>(ix.values);
            uint64_t acc = 0;
            for (int p = r0; p < r1; ++p)
                acc += static_cast<uint64_t>(static_cast<uint32_t>(vv[p]) *
                                             static_cast<uint32_t>(QL.value(ix.indices[p])));
            sc[j] = static_cast<float>(acc);
        } else {
            const float* vv = static_cast<const float*>(ix.values);
            float acc = 0.0F;
            for (int p = r0; p < r1; ++p) acc += vv[p] * QL.value(ix.indices[p]);
            sc[j] = acc;
        }
    }
    __syncthreads();
    float* out_d = a.out_d + q * a.k;
    int32_t* out_l = a.out_l + q * a.k;
    for (int j = tid; j < a.n_ids; j += kBlock) {
        const float s = sc[j];
        const int32_t id = a.ids[j];
        int rank = 0;
        for (int o = 0; o < a.n_ids; ++o) {
            const float os = sc[o];
            const int32_t oi = a.ids[o];
            if (better(os, oi, s, id) || (os == s && oi == id && o < j)) ++rank;
        }
        out_d[rank] = (SQ || W16) ? s * a.decode_scale : s;
        out_l[rank] = id;
    }
    for (int j = a.n_ids + tid; j < a.k; j += kBlock) {
        out_d[j] = -1.0F;
        out_l[j] = -1;
    }
}

// Exact FP32 scores of every document against queries, forward-index CSR
// (the scoring loop of BrutalIndex::single_query, brutal_index.cpp:75-96):
// the query sits in LDS as bitmap + rank table (a repeated term keeps its last
// value, as get_dense_vector_float does), one lane per document adds
// weight * query_value in row order, as the generic dot product does. Query
// q0 + blockIdx.y of the CSR batch (q_indptr) writes row blockIdx.y of
// `scores` (row stride ld); without q_indptr, blockIdx.y is 0 and the query is
// (q_indices, q_values, qn).
constexpr int kExactWords = 8;  // 8-entry words per step of a row in exact_scores_kernel

template <bool kVec>
__global__ __launch_bounds__(kBlock) void exact_scores_kernel(const int32_t* __restrict__ indptr,
                                                              const uint16_t* __restrict__ indices,
                                                              const float* __restrict__ values, int64_t n_docs,
                                                              int nw, const int32_t* __restrict__ q_indptr,
                                                              const uint16_t* __restrict__ q_indices,
                                                              const float* __restrict__ q_values, int qn,
                                                              int64_t q0, float* __restrict__ scores, int64_t ld) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* qbits = reinterpret_cast<uint32_t*>(smem);
    uint16_t* qpref = reinterpret_cast<uint16_t*>(qbits + nw);
    float* rval = reinterpret_cast<float*>(smem + ((nw * 6 + 15) & ~15));
    int32_t* rlast = reinterpret_cast<int32_t*>(rval + kQMax);
    int32_t* wsum = rlast + kQMax;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (q_indptr != nullptr) {
        const int64_t q = q0 + blockIdx.y;
        q_indices += q_indptr[q];
        q_values += q_indptr[q];
        qn = q_indptr[q + 1] - q_indptr[q];
        scores += static_cast<int64_t>(blockIdx.y) * ld;
    }
    for (int w = tid; w < nw; w += kBlock) qbits[w] = 0;
    for (int r = tid; r < qn; r += kBlock) rlast[r] = -1;
    __syncthreads();
    for (int i = tid; i < qn; i += kBlock) atomicOr(&qbits[q_indices[i] >> 5], 1u << (q_indices[i] & 31));
    __syncthreads();
    {
        const int per = (nw + kBlock - 1) / kBlock;
        const int w0 = tid * per;
        int local = 0;
        for (int w = w0; w < w0 + per && w < nw; ++w) local += __popc(qbits[w]);
        const int incl = wave_incl_scan(local, lane);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int run = incl - local;
        for (int w = 0; w < wave; ++w) run += wsum[w];
        for (int w = w0; w < w0 + per && w < nw; ++w) {
            qpref[w] = static_cast<uint16_t>(run);
            run += __popc(qbits[w]);
        }
    }
    __syncthreads();
    const QueryLds QL{qbits, qpref, rval};
    for (int i = tid; i < qn; i += kBlock) {
        bool p;
        atomicMax(&rlast[QL.rank_of(q_indices[i], p)], i);
    }
    __syncthreads();
    for (int r = tid; r < qn; r += kBlock)
        if (rlast[r] >= 0) rval[r] = q_values[rlast[r]];
    __syncthreads();
    // kVec: the row is read as aligned words of 8 entries (16 bytes of term
    // ids, 32 of values) instead of one 2- and one 4-byte load per entry, so a
    // wave's loads touch an eighth as many cache lines per entry; the entries
    // are still added in row order. A word past the last entry of the arrays
    // (indptr[n_docs]) is read entry by entry.
    const int64_t nnz_end = kVec ? indptr[n_docs] : 0;
    for (int64_t d = blockIdx.x * static_cast<int64_t>(kBlock) + tid; d < n_docs;
         d += static_cast<int64_t>(gridDim.x) * kBlock) {
        const int32_t r0 = indptr[d], r1 = indptr[d + 1];
        float acc = 0.0F;
        if constexpr (kVec) {
            // kExactWords words of 8 entries per step, their loads in flight together
            constexpr int kW = kExactWords;
            const uint4* iw = reinterpret_cast<const uint4*>(indices);
            const float4* vw = reinterpret_cast<const float4*>(values);
            const int64_t wend = (static_cast<int64_t>(r1) + 7) >> 3;
            for (int64_t w0 = r0 >> 3; w0 < wend; w0 += kW) {
                if ((w0 + kW) * 8 <= nnz_end) {
                    uint4 ti[kW];
                    float4 va[kW], vb[kW];
#pragma unroll
                    for (int u = 0; u < kW; ++u) {
                        ti[u] = iw[w0 + u];
                        va[u] = vw[2 * (w0 + u)];
                        vb[u] = vw[2 * (w0 + u) + 1];
                    }
#pragma unroll
                    for (int u = 0; u < kW; ++u) {
                        const uint32_t tt[4] = {ti[u].x, ti[u].y, ti[u].z, ti[u].w};
                        const float vv[8] = {va[u].x, va[u].y, va[u].z, va[u].w, vb[u].x, vb[u].y, vb[u].z, vb[u].w};
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const int64_t p = (w0 + u) * 8 + e;
                            if (p >= r0 && p < r1) {
                                const uint32_t t = (e & 1) ? (tt[e >> 1] >> 16) : (tt[e >> 1] & 0xFFFFu);
                                acc += vv[e] * QL.value(t);
                            }
                        }
                    }
                } else {
                    for (int64_t p = w0 * 8 > r0 ? w0 * 8 : r0; p < r1; ++p) acc += values[p] * QL.value(indices[p]);
                }
            }
        } else {
            for (int32_t p = r0; p < r1; ++p) acc += values[p] * QL.value(indices[p]);
        }
        scores[d] = acc;
    }
}

namespace {

// Rows can be read as 16-byte words of term ids and values when both arrays
// start on a 16-byte boundary.
bool words_aligned(const uint16_t* indices, const float* values) {
    return reinterpret_cast<uintptr_t>(indices) % 16 == 0 && reinterpret_cast<uintptr_t>(values) % 16 == 0;
}

int exact_scores_lds(int dim) { return (((dim + 31) / 32 * 6 + 15) & ~15) + kQMax * 8 + kWaves * 4; }

unsigned exact_scores_blocks(int64_t n_docs) {
    return static_cast<unsigned>(std::max<int64_t>(std::min<int64_t>((n_docs + kBlock - 1) / kBlock, 4096), 1));
}

// Top-k selection over rows of scores: key = (order-preserving bits of the
// score) << 32 | ~doc, so the k largest keys are the k best docs by (score
// desc, doc id asc), the order in which the reference's reject-when-not-better
// heap keeps them (earlier docs win ties at the k-th).
constexpr int kSelectSlots = 4096;  // LDS keys: the sorted top-k, then candidates
constexpr int kSelectPer = 16;      // keys a thread holds per round (loads in flight)
constexpr int kSelectRound = kBlock * kSelectPer;

__device__ __forceinline__ uint64_t score_key(float s, int64_t d) {
    uint32_t u = __float_as_uint(s);
    u ^= (u >> 31) != 0 ? 0xFFFFFFFFu : 0x80000000u;
    return (static_cast<uint64_t>(u) << 32) | (0xFFFFFFFFu - static_cast<uint32_t>(d));
}

__device__ __forceinline__ float key_score(uint64_t key) {
    uint32_t u = static_cast<uint32_t>(key >> 32);
    u ^= (u >> 31) != 0 ? 0x80000000u : 0xFFFFFFFFu;
    return __uint_as_float(u);
}

// Descending bitonic sort of keys[0, n) (n a power of two) by all kBlock threads.
__device__ void bitonic_desc(uint64_t* keys, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < n / 2; t += kBlock) {
                const int i = 2 * t - (t & (stride - 1));
                const int j = i + stride;
                const bool desc = (i & size) == 0;
                const uint64_t a = keys[i], b = keys[j];
                if ((a < b) == desc) {
                    keys[i] = b;
                    keys[j] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Keeps the kcap best of n candidate keys (key_of(i); 0 = not a candidate) in
// keys[0, kcap), best first: the candidates stream through a filter against
// the current k-th key, kSelectPer keys per thread per round held in
// registers; survivors collect in LDS behind the sorted top-kcap. When the
// candidate slots run out, the placed keys are merged in by a bitonic sort
// (which raises the filter) and the round's unplaced keys are filtered and
// placed again. A key at or below the k-th key can never enter the top-k.
template <class KeyOf>
__device__ void stream_topk(uint64_t* keys, int* cnt, int64_t n, int k, int kcap, KeyOf key_of) {
    const int cap = kSelectSlots - kcap;  // candidate slots
    uint64_t thresh = 0;
    if (k <= kBlock) {
        // A first threshold: every thread's best key of the first round is a
        // real candidate, so the k-th largest of those kBlock keys is at most
        // the final k-th; starting the filter just below it spares the early
        // merges of nearly every candidate.
        uint64_t best = 0;
#pragma unroll
        for (int u = 0; u < kSelectRound / kBlock; ++u) {
            const int64_t i = u * kBlock + threadIdx.x;
            if (i < n) {
                const uint64_t key = key_of(i);
                best = key > best ? key : best;
            }
        }
        keys[threadIdx.x] = best;
        __syncthreads();
        for (int size = 2; size <= kBlock; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const int t = threadIdx.x;
                if (t < kBlock / 2) {
                    const int i = 2 * t - (t & (stride - 1)), j = i + stride;
                    const bool desc = (i & size) == 0;
                    const uint64_t a = keys[i], b = keys[j];
                    if ((a < b) == desc) {
                        keys[i] = b;
                        keys[j] = a;
                    }
                }
                __syncthreads();
            }
        }
        const uint64_t kth = keys[k - 1];
        thresh = kth > 0 ? kth - 1 : 0;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < kSelectSlots; i += kBlock) keys[i] = 0;
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
    auto merge = [&]() {
        // only the used slots, padded with zero keys to a power of two, are sorted
        const int used = kcap + min(*cnt, cap);
        int n2 = kcap;
        while (n2 < used) n2 <<= 1;
        for (int i = used + threadIdx.x; i < n2; i += kBlock) keys[i] = 0;
        __syncthreads();
        bitonic_desc(keys, n2);
        if (threadIdx.x == 0) *cnt = 0;
        thresh = keys[k - 1];
        __syncthreads();
    };
    for (int64_t base = 0; base < n; base += kSelectRound) {
        uint64_t v[kSelectPer];
#pragma unroll
        for (int u = 0; u < kSelectPer; ++u) {
            const int64_t i = base + u * kBlock + threadIdx.x;
            v[u] = i < n ? key_of(i) : 0;
        }
        uint32_t pending = 0;  // bit u: v[u] passed the filter and has no slot yet
#pragma unroll
        for (int u = 0; u < kSelectPer; ++u)
            if (v[u] > thresh) pending |= 1u << u;
        while (true) {
#pragma unroll
            for (int u = 0; u < kSelectPer; ++u) {
                if (((pending >> u) & 1u) == 0) continue;
                if (v[u] > thresh) {
                    const int slot = atomicAdd(cnt, 1);
                    if (slot >= cap) continue;  // no room: placed after the next merge
                    keys[kcap + slot] = v[u];
                }
                pending &= ~(1u << u);
            }
            __syncthreads();
            const int c = *cnt;
            __syncthreads();  // every thread has read cnt before a merge resets it
            if (c < cap) break;  // every survivor has a slot
            merge();
        }
    }
    if (*cnt > 0) merge();
}

// Doc ids per top-k chunk: each row of scores is cut into chunks of this many
// docs, selected by workgroups of their own, then merged.
constexpr int64_t kSelectChunk = 65536;

int64_t select_chunks(int64_t n_docs) { return std::max<int64_t>(1, (n_docs + kSelectChunk - 1) / kSelectChunk); }

// Stage 1, one workgroup per (chunk, row): the k best docs of the chunk by
// (score desc, doc id asc) -- key = (order-preserving bits of the score) << 32
// | ~doc -- written as kcap keys (0 = none) to partial[row][chunk]. With
// `positive`, only scores > 0 are candidates; with `allow`, only docs whose
// bit is set.
__global__ __launch_bounds__(kBlock) void topk_chunk_kernel(const float* __restrict__ scores, int64_t ld,
                                                            int64_t n_docs, int k, int kcap, bool positive,
                                                            const uint32_t* __restrict__ allow,
                                                            uint64_t* __restrict__ partial) {
    __shared__ uint64_t keys[kSelectSlots];
    __shared__ int cnt;
    const float* row = scores + static_cast<int64_t>(blockIdx.y) * ld;
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * kSelectChunk;
    const int64_t n = min(kSelectChunk, n_docs - lo);
    stream_topk(keys, &cnt, n, k, kcap, [&](int64_t i) -> uint64_t {
        const int64_t d = lo + i;
        const float sc = row[d];
        bool keep = !positive || sc > 0.0F;
        if (allow != nullptr) keep = keep && ((allow[d >> 5] >> (d & 31)) & 1U) != 0;
        return keep ? score_key(sc, d) : 0;
    });
    uint64_t* out = partial + (static_cast<int64_t>(blockIdx.y) * gridDim.x + blockIdx.x) * kcap;
    for (int i = threadIdx.x; i < kcap; i += kBlock) out[i] = keys[i];
}

// Stage 2, one workgroup per row: the k best of the row's n_chunks * kcap
// partial keys, written as k (score, doc) pairs best first, (-1.0f, -1) past
// the last candidate.
__global__ __launch_bounds__(kBlock) void topk_merge_kernel(const uint64_t* __restrict__ partial, int n_chunks,
                                                            int k, int kcap, float* __restrict__ out_d,
                                                            int32_t* __restrict__ out_l) {
    __shared__ uint64_t keys[kSelectSlots];
    __shared__ int cnt;
    const uint64_t* in = partial + static_cast<int64_t>(blockIdx.x) * n_chunks * kcap;
    stream_topk(keys, &cnt, static_cast<int64_t>(n_chunks) * kcap, k, kcap, [&](int64_t i) { return in[i]; });
    float* od = out_d + static_cast<int64_t>(blockIdx.x) * k;
    int32_t* ol = out_l + static_cast<int64_t>(blockIdx.x) * k;
    for (int i = threadIdx.x; i < k; i += kBlock) {
        const uint64_t key = keys[i];
        od[i] = key != 0 ? key_score(key) : -1.0F;
        ol[i] = key != 0 ? static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(key)) : -1;
    }
}

// Both stages over g rows of scores; partial holds g * select_chunks(n_docs) * kcap keys.
void launch_topk(const float* scores, int64_t n_docs, int64_t g, int k, int kcap, bool positive,
                 const uint32_t* allow, uint64_t* partial, float* out_d, int32_t* out_l, hipStream_t stream) {
    const int64_t chunks = select_chunks(n_docs);
    hipLaunchKernelGGL(topk_chunk_kernel, dim3(static_cast<unsigned>(chunks), static_cast<unsigned>(g)),
                       dim3(kBlock), 0, stream, scores, n_docs, n_docs, k, kcap, positive, allow, partial);
    hipLaunchKernelGGL(topk_merge_kernel, dim3(static_cast<unsigned>(g)), dim3(kBlock), 0, stream, partial,
                       static_cast<int>(chunks), k, kcap, out_d, out_l);
}

}  // namespace

void launch_exact_scores(const int32_t* indptr, const uint16_t* indices, const float* values, int64_t n_docs, int dim,
                         const uint16_t* q_indices, const float* q_values, int qn, float* scores, hipStream_t stream) {
    if (qn > kQMax) throw std::invalid_argument("query has more non-zeros than the device path supports (512)");
    const int nw = (dim + 31) / 32;
    if (words_aligned(indices, values))
        hipLaunchKernelGGL(exact_scores_kernel<true>, dim3(exact_scores_blocks(n_docs)), dim3(kBlock),
                           exact_scores_lds(dim), stream, indptr, indices, values, n_docs, nw, nullptr, q_indices,
                           q_values, qn, int64_t{0}, scores, int64_t{0});
    else
        hipLaunchKernelGGL(exact_scores_kernel<false>, dim3(exact_scores_blocks(n_docs)), dim3(kBlock),
                           exact_scores_lds(dim), stream, indptr, indices, values, n_docs, nw, nullptr, q_indices,
                           q_values, qn, int64_t{0}, scores, int64_t{0});
    hip_check(hipGetLastError(), "exact scores launch");
}

// Per group of queries: their rows of n_docs float scores, then the top-k
// partial keys (select_chunks(n_docs) chunks of up to kBruteKMax keys a row).
size_t brute_search_scratch_bytes(int64_t n_docs, int64_t n_queries) {
    const size_t g = static_cast<size_t>(brute_search_group(n_docs, n_queries));
    return ((g * static_cast<size_t>(n_docs) * sizeof(float) + 255) & ~size_t{255}) +
           g * static_cast<size_t>(select_chunks(n_docs)) * kBruteKMax * sizeof(uint64_t);
}

namespace {
uint64_t* partial_keys(float* d_scores, int64_t n_docs, int64_t n_queries) {
    const size_t g = static_cast<size_t>(brute_search_group(n_docs, n_queries));
    const size_t off = (g * static_cast<size_t>(n_docs) * sizeof(float) + 255) & ~size_t{255};
    return reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(d_scores) + off);
}
}  // namespace

int64_t brute_search_group(int64_t n_docs, int64_t n_queries) {
    constexpr int64_t kScoreBytes = int64_t{256} << 20;  // score rows held at once
    int64_t per = std::max<int64_t>(1, kScoreBytes / std::max<int64_t>(n_docs * 4, 1));
    if (const char* e = std::getenv("NSPARSE_BRUTE_GROUP_QUERIES")) per = std::min<int64_t>(per, std::max(1, std::atoi(e)));
    return std::max<int64_t>(1, std::min<int64_t>({per, n_queries, 65535}));
}

void launch_brute_search(const int32_t* indptr, const uint16_t* indices, const float* values, int64_t n_docs, int dim,
                         const int32_t* d_q_indptr, const uint16_t* d_q_indices, const float* d_q_values,
                         int64_t n_queries, int k, float* d_out_d, int32_t* d_out_l, float* d_scores,
                         hipStream_t stream) {
    if (k <= 0 || k > kBruteKMax) throw std::invalid_argument("brutal search supports 1 <= k <= 1024 on the device");
    int kcap = 1;
    while (kcap < k) kcap <<= 1;
    const int64_t group = brute_search_group(n_docs, n_queries);
    const int nw = (dim + 31) / 32;
    for (int64_t q0 = 0; q0 < n_queries; q0 += group) {
        const int64_t g = std::min<int64_t>(group, n_queries - q0);
        if (words_aligned(indices, values))
            hipLaunchKernelGGL(exact_scores_kernel<true>, dim3(exact_scores_blocks(n_docs), static_cast<unsigned>(g)),
                               dim3(kBlock), exact_scores_lds(dim), stream, indptr, indices, values, n_docs, nw,
                               d_q_indptr, d_q_indices, d_q_values, 0, q0, d_scores, n_docs);
        else
            hipLaunchKernelGGL(exact_scores_kernel<false>, dim3(exact_scores_blocks(n_docs), static_cast<unsigned>(g)),
                               dim3(kBlock), exact_scores_lds(dim), stream, indptr, indices, values, n_docs, nw,
                               d_q_indptr, d_q_indices, d_q_values, 0, q0, d_scores, n_docs);
        launch_topk(d_scores, n_docs, g, k, kcap, false, nullptr, partial_keys(d_scores, n_docs, n_queries),
                    d_out_d + q0 * k, d_out_l + q0 * k, stream);
        hip_check(hipGetLastError(), "brute-force search launch");
    }
}

namespace {

// ---- InvertedIndex: exact term-at-a-time scores in LDS windows of doc ids
constexpr int kInvBlock = 1024;
constexpr int kInvWindow = kInvertedWindow;  // 64 KB of float scores
constexpr int kInvChunks = 4;                // a round takes up to kInvChunks * kInvBlock postings of one term
constexpr int kInvRound = kInvChunks * kInvBlock;
constexpr int kInvSlabTarget = 2048;         // workgroups a group of queries aims for

// window_start[t * (nw + 1) + w] = lower_bound(list t's doc ids, w * kInvWindow),
// one thread per (list, window boundary).
__global__ void window_start_kernel(const int64_t* __restrict__ list_ptr, const int32_t* __restrict__ docs, int dim,
                                    int64_t nw, int32_t* __restrict__ out) {
    const int64_t total = static_cast<int64_t>(dim) * (nw + 1);
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t t = i / (nw + 1), w = i % (nw + 1);
        const int64_t key = w * kInvWindow;
        int64_t a = list_ptr[t], b = list_ptr[t + 1];
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (docs[m] < key)
                a = m + 1;
            else
                b = m;
        }
        out[i] = static_cast<int32_t>(a);
    }
}

// Postings a round holds in registers: lane tid has positions
// p0 + j * kInvBlock + tid for j < kInvChunks below p0 + n; the others read as
// doc INT32_MAX (never inside a window).
struct InvRound {
    int32_t d[kInvChunks];
    float v[kInvChunks];
};

__device__ __forceinline__ void inv_load(const int32_t* __restrict__ docs, const float* __restrict__ vals, int64_t p0,
                                         int n, int tid, InvRound& r) {
#pragma unroll
    for (int j = 0; j < kInvChunks; ++j) {
        const int off = j * kInvBlock + tid;
        const bool ok = off < n;
        r.d[j] = ok ? docs[p0 + off] : INT32_MAX;
        r.v[j] = ok ? vals[p0 + off] : 0.0F;
    }
}

// One workgroup per (slab of consecutive windows, query). Each window is
// scored term by term in query order. The window_start table gives each term's
// postings inside the window exactly, so rounds of up to kInvRound postings
// load only those (no posting is read twice). Every posting adds weight * value
// into its doc's slot. A term's docs are distinct, so its rounds need no
// barrier between them; a barrier ends each term, so a doc's score is the
// float sum of its products in query-term order, from 0 -- the same bits
// however the work is split. The round after the current one (the term's
// next, or the next term's first) loads while the current one is added. The
// window's scores are then written to row blockIdx.y of `scores` (row stride
// ld). kAtomic: some list holds a doc twice, so one round may add to a slot
// twice.
template <bool kAtomic>
__global__ __launch_bounds__(kInvBlock) void inverted_scores_kernel(
    const int32_t* __restrict__ wstart, int64_t n_windows, const int32_t* __restrict__ docs,
    const float* __restrict__ vals, int64_t n_docs, int windows_per_slab, const int32_t* __restrict__ q_indptr,
    const uint16_t* __restrict__ q_indices, const float* __restrict__ q_values, int64_t q0,
    float* __restrict__ scores, int64_t ld) {
    __shared__ float win[kInvWindow];
    __shared__ int32_t lo[kQMax];  // list positions fit 32 bits: the host caps postings at 2^31 - 1
    __shared__ int32_t hi[kQMax];
    __shared__ int32_t term[kQMax];
    __shared__ float wt[kQMax];
    const int tid = threadIdx.x;
    const int64_t q = q0 + blockIdx.y;
    const int32_t qs = q_indptr[q];
    const int qn = q_indptr[q + 1] - qs;
    const int64_t w0 = static_cast<int64_t>(blockIdx.x) * windows_per_slab;
    const int64_t w1 = min(n_windows, w0 + windows_per_slab);
    float* row = scores + static_cast<int64_t>(blockIdx.y) * ld;
    for (int t = tid; t < qn; t += kInvBlock) {
        term[t] = q_indices[qs + t];
        wt[t] = q_values[qs + t];
    }
    for (int i = tid; i < kInvWindow; i += kInvBlock) win[i] = 0.0F;
    for (int64_t w = w0; w < w1; ++w) {
        const int64_t base = w * kInvWindow;
        const int64_t wend = min(base + kInvWindow, n_docs);
        __syncthreads();  // term[] / wt[] written, the previous window's lo / hi read
        for (int t = tid; t < qn; t += kInvBlock) {
            const int32_t* ws = wstart + static_cast<int64_t>(term[t]) * (n_windows + 1) + w;
            lo[t] = ws[0];
            hi[t] = ws[1];
        }
        __syncthreads();
        int t = 0;
        while (t < qn && lo[t] == hi[t]) ++t;  // terms with nothing in the window are skipped
        InvRound c;
        int64_t c_p0 = 0;
        int c_n = 0;
        if (t < qn) {
            c_p0 = lo[t];
            c_n = min(kInvRound, hi[t] - lo[t]);
            inv_load(docs, vals, c_p0, c_n, tid, c);
        }
        while (t < qn) {
            const int32_t tend = hi[t];
            const float wgt = wt[t];
            int t2 = t;
            int64_t p2 = c_p0 + c_n;
            if (p2 >= tend) {  // the term's last round: the next is the next term's first
                t2 = t + 1;
                while (t2 < qn && lo[t2] == hi[t2]) ++t2;
                p2 = t2 < qn ? lo[t2] : 0;
            }
            const int n2 = t2 < qn ? min(kInvRound, static_cast<int>((t2 == t ? tend : hi[t2]) - p2)) : 0;
            InvRound nx;
            if (t2 < qn) inv_load(docs, vals, p2, n2, tid, nx);
#pragma unroll
            for (int j = 0; j < kInvChunks; ++j) {
                if (c.d[j] != INT32_MAX) {
                    const float add = wgt * c.v[j];
                    float* slot = &win[c.d[j] - base];
                    if (kAtomic)
                        atomicAdd(slot, add);
                    else
                        *slot += add;
                }
            }
            if (t2 != t) __syncthreads();  // the next term adds to the same slots
            t = t2;
            c = nx;
            c_p0 = p2;
            c_n = n2;
        }
        __syncthreads();
        for (int i = tid; i < wend - base; i += kInvBlock) {
            row[base + i] = win[i];
            win[i] = 0.0F;
        }
    }
}

}  // namespace

void build_window_starts(DevicePostings& p, int64_t n_docs, int dim) {
    p.window_start.release();
    p.n_windows = 0;
    // positions are int32 (the search refuses more postings anyway)
    if (n_docs <= 0 || dim <= 0 || p.n_postings > std::numeric_limits<int32_t>::max()) return;
    const int64_t nw = (n_docs + kInvWindow - 1) / kInvWindow;
    p.window_start.allocate(static_cast<size_t>(dim) * static_cast<size_t>(nw + 1) * 4);
    p.n_windows = nw;
    const int64_t total = static_cast<int64_t>(dim) * (nw + 1);
    const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536)));
    hipLaunchKernelGGL(window_start_kernel, dim3(grid), dim3(256), 0, nullptr, p.list_ptr.get<int64_t>(),
                       p.docs.get<int32_t>(), dim, nw, p.window_start.get<int32_t>());
    hip_check(hipGetLastError(), "window start launch");
    hip_check(hipDeviceSynchronize(), "window starts");
}

void launch_inverted_search(const DevicePostings& ps, int64_t n_docs, const int32_t* d_q_indptr,
                            const uint16_t* d_q_indices, const float* d_q_values, int64_t n_queries, int k,
                            const uint32_t* d_allow, float* d_out_d, int32_t* d_out_l, float* d_scores,
                            hipStream_t stream) {
    if (k <= 0 || k > kBruteKMax) throw std::invalid_argument("inverted search supports 1 <= k <= 1024 on the device");
    if (n_docs <= 0 || n_queries <= 0) throw std::invalid_argument("inverted search needs documents and queries");
    if (ps.n_postings > std::numeric_limits<int32_t>::max())
        throw std::invalid_argument("inverted search addresses at most 2^31 - 1 postings");
    const int64_t windows = (n_docs + kInvWindow - 1) / kInvWindow;
    if (ps.n_windows != windows || !ps.window_start.bytes())
        throw std::logic_error("inverted search without the window table of its lists");
    int kcap = 1;
    while (kcap < k) kcap <<= 1;
    const int64_t group = brute_search_group(n_docs, n_queries);
    for (int64_t q0 = 0; q0 < n_queries; q0 += group) {
        const int64_t g = std::min<int64_t>(group, n_queries - q0);
        const int64_t wps = std::clamp<int64_t>(windows * g / kInvSlabTarget, 1, windows);
        const int64_t slabs = (windows + wps - 1) / wps;  // every slab starts below n_docs
        const dim3 grid(static_cast<unsigned>(slabs), static_cast<unsigned>(g));
        if (ps.duplicate_docs)
            hipLaunchKernelGGL(inverted_scores_kernel<true>, grid, dim3(kInvBlock), 0, stream,
                               ps.window_start.get<int32_t>(), windows, ps.docs.get<int32_t>(), ps.vals.get<float>(),
                               n_docs, static_cast<int>(wps), d_q_indptr, d_q_indices, d_q_values, q0, d_scores,
                               n_docs);
        else
            hipLaunchKernelGGL(inverted_scores_kernel<false>, grid, dim3(kInvBlock), 0, stream,
                               ps.window_start.get<int32_t>(), windows, ps.docs.get<int32_t>(), ps.vals.get<float>(),
                               n_docs, static_cast<int>(wps), d_q_indptr, d_q_indices, d_q_values, q0, d_scores,
                               n_docs);
        launch_topk(d_scores, n_docs, g, k, kcap, true, d_allow, partial_keys(d_scores, n_docs, n_queries),
                    d_out_d + q0 * k, d_out_l + q0 * k, stream);
        hip_check(hipGetLastError(), "inverted search launch");
    }
}

int search_lds_bytes(int dim, int max_clusters, int kmax, bool sq) {
    const SearchConfig c = resolve_search_config(dim, max_clusters, kmax, sq, default_search_config(sq));
    return make_plan(dim, max_clusters, kmax, sq, c.lookup == 1, c.block).total;
}

std::string device_error_message(int32_t code) {
    switch (code) {
        case kDevQueryTooLong: return "query has more non-zeros than the device path supports (512)";
        case kDevTermOutOfRange: return "query term id out of range";
        case kDevTooManyClusters: return "posting list has more clusters than the device plan";
        case kDevBadK: return "per-query k outside [1, k_max]";
        case kDevChunkWait: return "a wait between the waves of a search workgroup timed out";
        case kDevBadSelector: return "per-query selector index outside the bitmap table";
        default: return "device search reported error " + std::to_string(code);
    }
}

namespace {

// ---- longest-first order for the persistent schedule (all plain kernels, so a
// launch sequence can be captured in a graph: no memsets, no library calls)
constexpr int kOrderBuckets = kQMax + 1;

__global__ void zero_i32_kernel(int32_t* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// bucket b = kQMax - min(nnz, kQMax): bucket 0 holds the longest queries
__global__ void nnz_bucket_kernel(const int32_t* __restrict__ q_indptr, int64_t n, int32_t* __restrict__ hist) {
    for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int len = static_cast<int>(min<int64_t>(q_indptr[q + 1] - q_indptr[q], kQMax));
        atomicAdd(&hist[kQMax - max(len, 0)], 1);
    }
}

// exclusive scan of the bucket counts, in place (one wavefront)
__global__ void bucket_scan_kernel(int32_t* hist, int nb) {
    const int lane = threadIdx.x;
    int run = 0;
    for (int b0 = 0; b0 < nb; b0 += 64) {
        const int b = b0 + lane;
        const int v = b < nb ? hist[b] : 0;
        const int incl = wave_incl_scan(v, lane);
        if (b < nb) hist[b] = run + incl - v;
        run += __shfl(incl, 63, 64);
    }
}

__global__ void order_scatter_kernel(const int32_t* __restrict__ q_indptr, int64_t n, int32_t* __restrict__ offs,
                                     int32_t* __restrict__ order) {
    for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int len = static_cast<int>(min<int64_t>(q_indptr[q + 1] - q_indptr[q], kQMax));
        order[atomicAdd(&offs[kQMax - max(len, 0)], 1)] = static_cast<int32_t>(q);
    }
}

int device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    return cus;
}

template <bool SQ, bool DENSE, int BLOCK, bool W16 = false>
void launch_variant(const SearchArgs& args, const LdsPlan& P, bool persistent, hipStream_t stream) {
    const void* kfn = persistent
                          ? reinterpret_cast<const void*>(seismic_search_persistent_kernel<SQ, DENSE, BLOCK, W16>)
                          : reinterpret_cast<const void*>(seismic_search_kernel<SQ, DENSE, BLOCK, W16>);
    if (P.total > 64 * 1024)
        hip_check(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, P.total),
                  "hipFuncSetAttribute");
    if (persistent) {
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, BLOCK, P.total), "occupancy");
        const int64_t want = static_cast<int64_t>(std::max(per_cu, 1)) * device_cus();
        const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(want, args.n_queries)));
        hipLaunchKernelGGL((seismic_search_persistent_kernel<SQ, DENSE, BLOCK, W16>), dim3(grid), dim3(BLOCK),
                           P.total, stream, args, P);
        hip_check(hipGetLastError(), "search kernel launch");
        return;
    }
    constexpr int64_t kMaxGrid = 1 << 30;
    for (int64_t base = 0; base < args.n_queries; base += kMaxGrid) {
        SearchArgs a = args;
        a.query_base = base;
        const int64_t nb = args.n_queries - base < kMaxGrid ? args.n_queries - base : kMaxGrid;
        hipLaunchKernelGGL((seismic_search_kernel<SQ, DENSE, BLOCK, W16>), dim3(static_cast<unsigned>(nb)),
                           dim3(BLOCK), P.total, stream, a, P);
        hip_check(hipGetLastError(), "search kernel launch");
    }
}

}  // namespace

SearchConfig default_search_config(bool sq) {
    SearchConfig c{};
    c.lookup = sq ? 1 : 0;
    const char* e = std::getenv("NSPARSE_SEARCH_LOOKUP");
    if (sq && e != nullptr && std::string(e) == "bitmap") c.lookup = 0;
    c.block = sq && c.lookup == 1 ? 512 : 256;
    const char* eb = std::getenv("NSPARSE_SEARCH_BLOCK");
    if (sq && eb != nullptr) c.block = std::string(eb) == "1024" ? 1024 : std::string(eb) == "512" ? 512 : 256;
    c.persistent = 0;
    c.longest_first = 0;
    return c;
}

SearchConfig resolve_search_config(int dim, int max_clusters, int kmax, bool sq, SearchConfig c) {
    if (!sq) {
        c.lookup = 0;
        c.block = 256;
    }
    if (c.block != 256 && c.block != 512 && c.block != 1024)
        throw std::invalid_argument("search workgroup size must be 256, 512 or 1024");
    constexpr int kLds = 160 * 1024;
    // 1024 lanes only with the dense table, and only when its plan fits
    if (c.block == 1024 && (c.lookup != 1 || make_plan(dim, max_clusters, kmax, sq, true, 1024).total > kLds))
        c.block = 512;
    if (c.lookup == 1 && make_plan(dim, max_clusters, kmax, sq, true, c.block).total > kLds) c.lookup = 0;
    if (c.lookup == 0 && c.block == 1024) c.block = 512;
    if (make_plan(dim, max_clusters, kmax, sq, c.lookup == 1, c.block).total > kLds)
        throw std::invalid_argument("search does not fit in LDS for this dimension / cluster count / k");
    if (!c.persistent) c.longest_first = 0;
    return c;
}

size_t search_scratch_bytes(int64_t n) { return 16 + 4 * static_cast<size_t>((kOrderBuckets + 15) & ~15) + 4 * n; }

void launch_search(const SearchArgs& args_in, int max_clusters, int kmax, bool sq, const SearchConfig& cfg_in,
                   void* scratch, hipStream_t stream) {
    const SearchConfig cfg = resolve_search_config(args_in.ix.dim, max_clusters, kmax, sq, cfg_in);
    const LdsPlan P = make_plan(args_in.ix.dim, max_clusters, kmax, sq, cfg.lookup == 1, cfg.block);
    SearchArgs a = args_in;
    a.queue = nullptr;
    a.order = nullptr;
    if (const char* e = std::getenv("NSPARSE_SEARCH_SPIN_LIMIT")) a.spin_max = std::max(1, std::atoi(e));
    // the error flag reports this launch (a kernel, not a memset: graph-safe)
    // kernels read the flag too (a timed-out wait ends the launch's other waits)
    if (a.error_flag == nullptr) throw std::invalid_argument("search launch without an error flag");
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, stream, a.error_flag, 1);
    if (cfg.persistent) {
        if (scratch == nullptr) throw std::invalid_argument("the persistent schedule needs a scratch buffer");
        int32_t* base = static_cast<int32_t*>(scratch);
        int32_t* hist = base + 4;
        int32_t* order = hist + ((kOrderBuckets + 15) & ~15);
        a.queue = base;
        const int nz = cfg.longest_first ? 4 + kOrderBuckets : 1;
        hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(256), 0, stream, base, nz);
        if (cfg.longest_first) {
            const unsigned g = static_cast<unsigned>(std::min<int64_t>((a.n_queries + 255) / 256, 4096));
            hipLaunchKernelGGL(nnz_bucket_kernel, dim3(g), dim3(256), 0, stream, a.q_indptr, a.n_queries, hist);
            hipLaunchKernelGGL(bucket_scan_kernel, dim3(1), dim3(64), 0, stream, hist, kOrderBuckets);
            hipLaunchKernelGGL(order_scatter_kernel, dim3(g), dim3(256), 0, stream, a.q_indptr, a.n_queries, hist,
                               order);
            a.order = order;
        }
        hip_check(hipGetLastError(), "search queue launch");
    }
    const bool pers = cfg.persistent != 0;
    if (!sq && a.ix.esz == 2)
        launch_variant<false, false, 256, true>(a, P, pers, stream);
    else if (!sq)
        launch_variant<false, false, 256>(a, P, pers, stream);
    else if (cfg.lookup == 1 && cfg.block == 1024)
        launch_variant<true, true, 1024>(a, P, pers, stream);
    else if (cfg.lookup == 1 && cfg.block == 512)
        launch_variant<true, true, 512>(a, P, pers, stream);
    else if (cfg.lookup == 1)
        launch_variant<true, true, 256>(a, P, pers, stream);
    else if (cfg.block == 512)
        launch_variant<true, false, 512>(a, P, pers, stream);
    else
        launch_variant<true, false, 256>(a, P, pers, stream);
}

void launch_search_checked(const SearchArgs& args, int max_clusters, int kmax, bool sq, hipStream_t stream) {
    SearchConfig c = default_search_config(sq);
    // A launch of few queries (the per-query drop-in, small serving batches)
    // leaves most CUs idle: one 1024-lane workgroup per query, alone on its
    // CU, finishes it sooner than a 512-lane one. NSPARSE_SEARCH_BLOCK wins.
    if (sq && c.lookup == 1 && std::getenv("NSPARSE_SEARCH_BLOCK") == nullptr &&
        args.n_queries <= std::max(1, device_cus() / 4))
        c.block = 1024;
    launch_search(args, max_clusters, kmax, sq, c, nullptr, stream);
}

namespace {

// Scratch of the allocation-free batched search: the persistent schedule's
// queue counter, its nnz-bucket histogram and the longest-first query order,
// in the layout launch_search gives a caller's scratch buffer. Every GPU the
// library runs on holds its own copy until the process ends.
constexpr int64_t kNoAllocWords = static_cast<int64_t>((16 + 4 * ((kOrderBuckets + 15) & ~15)) / 4) +
                                  kNoAllocMaxQueries;
static __device__ __align__(256) int32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (the array's name is
// only a host-side stand-in), resolved once per device.
void* scratch_on_current_device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (p == nullptr) throw device_error("hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

void launch_search_noalloc(const SearchArgs& a, int max_clusters, int kmax, bool sq, hipStream_t stream) {
    if (a.n_queries > kNoAllocMaxQueries)
        throw std::invalid_argument("the allocation-free search takes at most " + std::to_string(kNoAllocMaxQueries) +
                                    " queries per call");
    static_assert(static_cast<size_t>(kNoAllocWords) * 4 >= 16 + 4 * ((kOrderBuckets + 15) & ~15) +
                                                                  4 * static_cast<size_t>(kNoAllocMaxQueries),
                  "scratch holds the persistent schedule's queue, histogram and order");
    void* ptr = scratch_on_current_device();
    SearchConfig c = default_search_config(sq);
    c.persistent = 1;
    c.longest_first = 1;
    launch_search(a, max_clusters, kmax, sq, c, ptr, stream);
}

void launch_exact_checked(const ExactArgs& a, bool sq, hipStream_t stream) {
    if (a.n_ids > kQMax) throw std::invalid_argument("exact-match selector larger than the device path supports");
    const int nw = (a.ix.dim + 31) / 32;
    const int lds = ((nw * 6 + 15) & ~15) + kQMax * 12 + kWaves * 4;
    const dim3 grid(static_cast<unsigned>(a.n_queries));
    if (sq)
        hipLaunchKernelGGL((exact_match_kernel<true>), grid, dim3(kBlock), lds, stream, a, nw);
    else if (a.ix.esz == 2)
        hipLaunchKernelGGL((exact_match_kernel<false, true>), grid, dim3(kBlock), lds, stream, a, nw);
    else
        hipLaunchKernelGGL((exact_match_kernel<false>), grid, dim3(kBlock), lds, stream, a, nw);
    hip_check(hipGetLastError(), "exact-match kernel launch");
}

}  // namespace nsparse_amd
