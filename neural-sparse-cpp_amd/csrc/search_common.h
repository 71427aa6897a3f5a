// Device helpers shared by the search kernels (search.hip, search_spread.hip):
// wave-level scans and ballots, the query's LDS tables, the step-exact
// std::sort emulation for the first list's visiting order, and the top-k merge.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_index.h"
#include "stl_order.h"

namespace nsparse_amd {

constexpr int kQMax = 512;  // longest query handled on device
constexpr int kMaxWaves = 16;

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Inclusive prefix sum over the 64 lanes with DPP moves only (no LDS
// crossbar): Hillis-Steele inside each 16-lane row (row_shr 1, 2, 4, 8), then
// row_bcast:15 and row_bcast:31 carry the row totals upward. Lanes whose DPP
// source is outside the row or masked off read the `old` operand, 0.
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false));
    return v;
}

typedef unsigned short ushort2_t __attribute__((ext_vector_type(2)));

// Base term of a block-row word (the layout in device_index.h): 0 for a
// low-term word (slot 0's code nonzero), else slot 0 itself.
__device__ __forceinline__ uint32_t blk_word_base(uint32_t w0) { return (w0 & 15U) != 0 ? 0U : (w0 & 0xFFF0U); }

// acc + the integer dot product of one block-row word with the query's
// vocabulary-sized code table: all 8 lookups first, then one dot2 per slot
// pair (each dword holds two slots, codes at bits 0-3 and 16-19). Slot 0 of a
// based word has code 0, so its lookup (of term base / 16) adds nothing.
__device__ __forceinline__ uint32_t blk_word_dot(const uint4& w, const uint8_t* qcode, uint32_t acc) {
    const uint32_t tw[4] = {w.x, w.y, w.z, w.w};
    const uint32_t base = blk_word_base(tw[0]);
    ushort2_t qp[4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        qp[h].x = qcode[(h == 0 ? 0U : base) + ((tw[h] >> 4) & 0xFFFU)];
        qp[h].y = qcode[base + (tw[h] >> 20)];
    }
#pragma unroll
    for (int h = 0; h < 4; ++h)
        acc = __builtin_amdgcn_udot2(__builtin_bit_cast(ushort2_t, tw[h] & 0x000F000FU), qp[h], acc, false);
    return acc;
}

__device__ __forceinline__ bool better(float s, int32_t i, float ts, int32_t ti) {
    return s > ts || (s == ts && i < ti);
}

struct QueryLds {
    const uint32_t* qbits;
    const uint16_t* qpref;
    const float* rval;
    __device__ __forceinline__ int rank_of(uint32_t t, bool& present) const {
        const uint32_t w = qbits[t >> 5];
        const uint32_t bit = 1u << (t & 31);
        present = (w & bit) != 0;
        return qpref[t >> 5] + __popc(w & (bit - 1));
    }
    __device__ __forceinline__ float value(uint32_t t) const {
        bool p;
        const int r = rank_of(t, p);
        return p ? rval[r] : 0.0F;
    }
};

// std::sort(A, A + n, greater_key) by one wavefront, step-exact with the
// libstdc++ introsort that stl_order.h restates for one lane (median-of-3
// pivot, unguarded Hoare partition, depth limit 2*floor(log2 n) with heap-sort
// fallback, 16-element threshold, final insertion sort).
//
// One Hoare partition of [first, last) with pivot p = A[first] (after the
// median move) is computed in parallel: with L_0 < L_1 < ... the positions in
// (first, last) whose key is <= p's ("left stoppers") and R_0 > R_1 > ... those
// whose key is >= p's ("right stoppers"), the sequential scan swaps
// A[L_i] <-> A[R_i] exactly while L_i < R_i (a swapped element stops the other
// scanner no later than the next original stopper), and returns
// min(L_m, R_{m-1}) after m swaps (L_0 when m = 0). The swaps are disjoint.
// Ranges are pushed on a lane-held stack (lane sp holds entry sp); the order in
// which disjoint ranges are finished does not change the result. Lpos and Rasc
// hold n positions each.
//
// The final insertion sort is left to the caller: after the partitions every
// key of an earlier leaf range (<= 16 elements, or a heap-sorted range) is >=
// every key of a later one, so __final_insertion_sort never moves an element
// out of its leaf and amounts to a stable sort of each leaf. `leaf_bits` (n
// bits) marks the first position of every leaf; heap-sorted ranges, already in
// order, mark every position.
inline __device__ void introsort_desc_wave(KeyItem* A, int n, uint16_t* Lpos, uint16_t* Rasc,
                                           uint32_t* leaf_bits, int lane) {
    for (int w = lane; w < (n + 31) / 32; w += 64) leaf_bits[w] = w == 0 ? 1u : 0u;
    __builtin_amdgcn_wave_barrier();
    if (n <= 1) return;
    const uint64_t lt = (1ULL << lane) - 1;
    int s_lo = 0, s_hi = n, s_d = floor_log2(n) * 2;  // entry `lane` of the stack
    int sp = 1;
    while (sp > 0) {
        --sp;
        int lo = __builtin_amdgcn_readlane(s_lo, sp);
        int hi = __builtin_amdgcn_readlane(s_hi, sp);
        int depth = __builtin_amdgcn_readlane(s_d, sp);
        bool heap = false;
        while (hi - lo > kIntrosortThreshold) {
            if (depth == 0) {
                if (lane == 0) {
                    heap_sort_range(A, lo, hi);
                    for (int x = lo; x < hi; ++x) leaf_bits[x >> 5] |= 1u << (x & 31);
                }
                __builtin_amdgcn_wave_barrier();
                heap = true;
                break;
            }
            --depth;
            if (lane == 0) median_to_first(A, lo, lo + 1, lo + (hi - lo) / 2, hi - 1);
            __builtin_amdgcn_wave_barrier();
            const float pk = A[lo].key;
            int nl = 0, nr = 0;
            for (int x0 = lo + 1; x0 < hi; x0 += 64) {
                const int x = x0 + lane;
                const bool in = x < hi;
                const float key = in ? A[x].key : 0.0F;
                const bool isl = in && !(key > pk);
                const bool isr = in && !(pk > key);
                const uint64_t ml = ballot(isl), mr = ballot(isr);
                if (isl) Lpos[nl + __popcll(ml & lt)] = static_cast<uint16_t>(x);
                if (isr) Rasc[nr + __popcll(mr & lt)] = static_cast<uint16_t>(x);
                nl += __popcll(ml);
                nr += __popcll(mr);
            }
            __builtin_amdgcn_wave_barrier();
            // R_i = Rasc[nr - 1 - i]; L_i < R_i holds for a prefix i < m
            const int pairs = min(nl, nr);
            int m = 0;
            for (int i0 = 0; i0 < pairs; i0 += 64) {
                const int i = i0 + lane;
                const bool c = i < pairs && Lpos[i] < Rasc[nr - 1 - i];
                m += __popcll(ballot(c));
            }
            for (int i0 = 0; i0 < m; i0 += 64) {
                const int i = i0 + lane;
                KeyItem a{}, b{};
                int pa = 0, pb = 0;
                if (i < m) {
                    pa = Lpos[i];
                    pb = Rasc[nr - 1 - i];
                    a = A[pa];
                    b = A[pb];
                }
                __builtin_amdgcn_wave_barrier();
                if (i < m) {
                    A[pa] = b;
                    A[pb] = a;
                }
            }
            int cut = m < nl ? static_cast<int>(Lpos[m]) : 0x7FFFFFFF;
            if (m > 0) cut = min(cut, static_cast<int>(Rasc[nr - m]));
            __builtin_amdgcn_wave_barrier();
            // push the right part, continue with the left one
            if (lane == sp) {
                s_lo = cut;
                s_hi = hi;
                s_d = depth;
            }
            ++sp;
            hi = cut;
        }
        if (!heap && lane == 0) leaf_bits[lo >> 5] |= 1u << (lo & 31);
        __builtin_amdgcn_wave_barrier();
    }
}

// introsort_desc_wave by a whole workgroup, with the same result: the
// partitions of one recursion level touch disjoint ranges, so each level's
// ranges are shared out over the waves, one level after another (the order in
// which disjoint ranges are finished does not change the result). A range
// [lo, hi) uses positions lo .. hi - 1 of Lpos and Rasc (n each). `rng` holds
// 2 * cap ranges (a level has at most 2 * n / 17 + 1 of them) and `cnt` two
// ints of LDS. Barriers inside; leaf_bits as for introsort_desc_wave.
template <int BLOCK>
__device__ void introsort_desc_block(KeyItem* A, int n, uint16_t* Lpos, uint16_t* Rasc, uint32_t* leaf_bits, int4* rng,
                                     int cap, int32_t* cnt) {
    constexpr int kW = BLOCK / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lt = (1ULL << lane) - 1;
    for (int w = tid; w < (n + 31) / 32; w += BLOCK) leaf_bits[w] = w == 0 ? 1u : 0u;
    if (tid == 0) {
        cnt[0] = n > 1 ? 1 : 0;
        cnt[1] = 0;
        rng[0] = make_int4(0, n, floor_log2(max(n, 1)) * 2, 0);
    }
    __syncthreads();
    for (int cur = 0;; cur ^= 1) {
        const int ncur = cnt[cur];
        if (ncur == 0) break;
        for (int r = wave; r < ncur; r += kW) {
            const int4 R = rng[cur * cap + r];
            const int lo = R.x, hi = R.y;
            int depth = R.z;
            if (hi - lo <= kIntrosortThreshold) {
                if (lane == 0 && lo < n) atomicOr(&leaf_bits[lo >> 5], 1u << (lo & 31));
                continue;
            }
            if (depth == 0) {
                if (lane == 0) {
                    heap_sort_range(A, lo, hi);
                    for (int x = lo; x < hi; ++x) atomicOr(&leaf_bits[x >> 5], 1u << (x & 31));
                }
                __builtin_amdgcn_wave_barrier();
                continue;
            }
            --depth;
            if (lane == 0) median_to_first(A, lo, lo + 1, lo + (hi - lo) / 2, hi - 1);
            __builtin_amdgcn_wave_barrier();
            uint16_t* L = Lpos + lo;
            uint16_t* Ra = Rasc + lo;
            const float pk = A[lo].key;
            int nl = 0, nr = 0;
            for (int x0 = lo + 1; x0 < hi; x0 += 64) {
                const int x = x0 + lane;
                const bool in = x < hi;
                const float key = in ? A[x].key : 0.0F;
                const bool isl = in && !(key > pk);
                const bool isr = in && !(pk > key);
                const uint64_t ml = ballot(isl), mr = ballot(isr);
                if (isl) L[nl + __popcll(ml & lt)] = static_cast<uint16_t>(x);
                if (isr) Ra[nr + __popcll(mr & lt)] = static_cast<uint16_t>(x);
                nl += __popcll(ml);
                nr += __popcll(mr);
            }
            __builtin_amdgcn_wave_barrier();
            const int pairs = min(nl, nr);
            int m = 0;
            for (int i0 = 0; i0 < pairs; i0 += 64) {
                const int i = i0 + lane;
                const bool c = i < pairs && L[i] < Ra[nr - 1 - i];
                m += __popcll(ballot(c));
            }
            for (int i0 = 0; i0 < m; i0 += 64) {
                const int i = i0 + lane;
                KeyItem va{}, vb{};
                int pa = 0, pb = 0;
                if (i < m) {
                    pa = L[i];
                    pb = Ra[nr - 1 - i];
                    va = A[pa];
                    vb = A[pb];
                }
                __builtin_amdgcn_wave_barrier();
                if (i < m) {
                    A[pa] = vb;
                    A[pb] = va;
                }
            }
            int cut = m < nl ? static_cast<int>(L[m]) : 0x7FFFFFFF;
            if (m > 0) cut = min(cut, static_cast<int>(Ra[nr - m]));
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                const int at = atomicAdd(&cnt[cur ^ 1], 2);
                rng[(cur ^ 1) * cap + at] = make_int4(lo, cut, depth, 0);
                rng[(cur ^ 1) * cap + at + 1] = make_int4(cut, hi, depth, 0);
            }
        }
        __syncthreads();
        if (tid == 0) cnt[cur] = 0;
        __syncthreads();
    }
}

// The final insertion sort as a stable sort of each leaf (see above): element
// i goes to leaf_start + (elements of its leaf with a greater key, or an equal
// key before it). Writes out[pos] = A[i].id. Any number of lanes.
inline __device__ void introsort_final_pass(const KeyItem* A, int n, const uint32_t* leaf_bits, uint16_t* out,
                                            int tid, int nthreads) {
    for (int i = tid; i < n; i += nthreads) {
        int w = i >> 5;
        uint32_t m = leaf_bits[w] & ((2u << (i & 31)) - 1u);  // marks at or before i
        while (m == 0) m = leaf_bits[--w];
        const int start = w * 32 + 31 - __clz(m);
        int end = n;
        const int i1 = i + 1;
        if (i1 < n) {
            w = i1 >> 5;
            m = leaf_bits[w] & ~((1u << (i1 & 31)) - 1u);  // marks at or after i + 1
            while (m == 0 && (w + 1) * 32 < n) m = leaf_bits[++w];
            if (m != 0) end = min(n, w * 32 + __ffs(m) - 1);
        }
        const float ki = A[i].key;
        int r = 0;
        for (int j = start; j < end; ++j) {
            const float kj = A[j].key;
            r += (kj > ki || (kj == ki && j < i)) ? 1 : 0;
        }
        out[start + r] = static_cast<uint16_t>(A[i].id);
    }
}

// A heap of at most 64 KeyItems in one wavefront's registers: element j sits
// in lane j, read with readlane and written by a lane-select at a
// wave-uniform index, so every lane runs the heap code in step.
struct LaneHeap {
    float key = 0.0F;
    uint32_t id = 0;
    int lane = 0;
};
struct LaneSlot {
    LaneHeap* h;
    int x;
    __device__ __forceinline__ operator KeyItem() const {
        return KeyItem{__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, h->key), x)),
                       static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(h->id), x))};
    }
    __device__ __forceinline__ LaneSlot& operator=(const KeyItem& v) {
        const bool mine = h->lane == x;
        h->key = mine ? v.key : h->key;
        h->id = mine ? v.id : h->id;
        return *this;
    }
    __device__ __forceinline__ LaneSlot& operator=(const LaneSlot& o) { return *this = static_cast<KeyItem>(o); }
};
struct LaneHeapRef {
    LaneHeap* h;
    __device__ __forceinline__ LaneSlot operator[](int x) const { return LaneSlot{h, x}; }
};

// topk_tokens (stl_order.h) by one wavefront, k <= 64, with its heap in
// registers and the same heap steps. Entries are taken 64 at a time; only
// those a ballot finds above the heap's threshold at the start of the 64 (all
// of them while the heap fills) go through the sequential steps, where each is
// tested again: the threshold only rises, so the others would be rejected by
// the sequential loop too. Lane 0 writes `out`; returns the count.
__device__ __forceinline__ int topk_tokens_wave(int n, int k, const float* weight, const uint16_t* term,
                                                uint16_t* out, int lane) {
    if (k >= n) {
        for (int i = lane; i < n; i += 64) out[i] = term[i];
        return n;
    }
    LaneHeap h;
    h.lane = lane;
    const LaneHeapRef heap{&h};
    int size = 0;
    float threshold = __builtin_inff();
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int i = c0 + lane;
        const bool in = i < n;
        const float s = in ? weight[i] : 0.0F;
        const int t = in ? static_cast<int>(term[i]) : 0;
        uint64_t cand = ballot(in && !(size >= k && s <= threshold));
        while (cand != 0) {
            const int l = __ffsll(static_cast<unsigned long long>(cand)) - 1;
            cand &= cand - 1;
            const float sv = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, s), l));
            const uint32_t tv = static_cast<uint32_t>(__builtin_amdgcn_readlane(t, l));
            if (size >= k && sv <= threshold) continue;
            if (size < k) {
                pq_push(heap, size, KeyItem{sv, tv});
                if (size == k) threshold = KeyItem(heap[0]).key;
            } else {
                pq_pop(heap, size);
                pq_push(heap, size, KeyItem{sv, tv});
                threshold = KeyItem(heap[0]).key;
            }
        }
    }
    const int total = size;
    for (int idx = total - 1; idx >= 0; --idx) {
        const uint32_t top = KeyItem(heap[0]).id;
        if (lane == 0) out[idx] = static_cast<uint16_t>(top);
        pq_pop(heap, size);
    }
    return total;
}

// Merges one wave's candidates (lane l: `in`, score `sc`, doc `id`) into the
// top-k: tks/tki hold tkn entries sorted by (score desc, doc id asc), at most
// k. Inserting the candidates one by one with the reference's reject rule keeps
// the k best of (top-k U candidates) under that order -- an order-free result
// -- so they are merged in parallel: each candidate's rank among the top-k
// (ballots against the top-k held in registers, or a binary search when it has
// more than 128 entries; an equal id is the same doc seen in an earlier list
// and is not inserted again) plus its rank among the other candidates. The
// candidates must be distinct docs. Uniform over the wave; tkn is updated.
// With more than kMergeSortMin candidates the wave sorts them first (a
// bitonic network over the 64 lanes on (score, doc) keys): a candidate's rank
// among the others is then a prefix count, its rank in the top-k a binary
// search of the top-k, and how far a top-k entry moves up a binary search over
// the lanes -- instead of one step per candidate in each of those counts.
constexpr int kMergeSortMin = 8;

// (score desc, doc id asc) as one unsigned order: order-preserving float bits
// above ~id. Never 0 for a doc id >= 0.
__device__ __forceinline__ uint64_t topk_order_key(float s, int32_t id) {
    if (s == 0.0F) s = 0.0F;  // -0 and +0 compare equal in better()
    uint32_t u = __float_as_uint(s);
    u ^= (u >> 31) != 0 ? 0xFFFFFFFFu : 0x80000000u;
    return (static_cast<uint64_t>(u) << 32) | (0xFFFFFFFFu - static_cast<uint32_t>(id));
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(v)), src, 64));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(static_cast<uint32_t>(v >> 32)), src, 64));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// kSortMany: take the sorted path for large batches (the spread search's
// replay, where a query's merges are its critical path; in the batched kernel
// the larger code measured 2.4% slower).
template <bool kSortMany = false>
__device__ __forceinline__ void merge_topk_wave(bool in, float sc, int32_t id, int k, int& tkn, float* tks,
                                                int32_t* tki, int lane) {
    const bool fullk = tkn == k;
    const float ks = fullk ? tks[k - 1] : 0.0F;
    const int32_t ki = fullk ? tki[k - 1] : 0;
    const bool cand = in && (!fullk || better(sc, id, ks, ki));
    const uint64_t cm = ballot(cand);
    if (cm == 0) return;
    if (kSortMany && __popcll(cm) > kMergeSortMin) {
        // lane i holds the i-th best candidate; key 0 past the last one
        uint64_t key = cand ? topk_order_key(sc, id) : 0;
        for (int size = 2; size <= 64; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const uint64_t o = shfl_u64(key, lane ^ stride);
                const bool keep_max = ((lane & stride) == 0) == ((lane & size) == 0);
                key = keep_max ? (key > o ? key : o) : (key < o ? key : o);
            }
        }
        const bool valid = key != 0;
        uint32_t u = static_cast<uint32_t>(key >> 32);
        u ^= (u >> 31) != 0 ? 0x80000000u : 0xFFFFFFFFu;
        const float csc = __uint_as_float(u);
        const int32_t cid = static_cast<int32_t>(0xFFFFFFFFu - static_cast<uint32_t>(key));
        int rt = 0;
        bool dup = false;
        if (valid) {
            int hi = tkn;
            while (rt < hi) {
                const int mid = (rt + hi) >> 1;
                if (better(tks[mid], tki[mid], csc, cid))
                    rt = mid + 1;
                else
                    hi = mid;
            }
            dup = rt < tkn && tki[rt] == cid;
        }
        const bool v = valid && !dup;
        const uint64_t vm = ballot(v);
        if (vm == 0) return;
        const int rc = __popcll(vm & ((1ULL << lane) - 1));
        // top-k entries move up past the kept candidates better than them,
        // highest tile first (a tile is read whole before it is written)
        for (int t0 = ((tkn - 1) >> 6) << 6; t0 >= 0 && tkn > 0; t0 -= 64) {
            const int t = t0 + lane;
            const bool live = t < tkn;
            const float es = live ? tks[t] : 0.0F;
            const int32_t ei = live ? tki[t] : 0;
            const uint64_t ek = live ? topk_order_key(es, ei) : ~0ULL;
            // candidates (descending over the lanes) with a key above ek: a prefix
            int pos = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1) {
                const uint64_t kk = shfl_u64(key, pos + step - 1);
                if (kk > ek) pos += step;
            }
            if (shfl_u64(key, pos < 64 ? pos : 63) > ek && pos < 64) ++pos;
            const int up = __popcll(vm & (pos >= 64 ? ~0ULL : ((1ULL << pos) - 1)));
            __builtin_amdgcn_wave_barrier();
            if (live && up > 0 && t + up < k) {
                tks[t + up] = es;
                tki[t + up] = ei;
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (v && rt + rc < k) {
            tks[rt + rc] = csc;
            tki[rt + rc] = cid;
        }
        __builtin_amdgcn_wave_barrier();
        tkn = min(k, tkn + __popcll(vm));
        return;
    }
    const auto lane_f = [](float v, int l) {
        return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    };
    int rt = 0;
    bool dup = false;
    const bool small = tkn <= 128;
    const bool l0 = lane < tkn, l1 = lane + 64 < tkn;
    float e0s = 0.0F, e1s = 0.0F;
    int32_t e0i = 0, e1i = 0;
    if (small) {
        if (l0) {
            e0s = tks[lane];
            e0i = tki[lane];
        }
        if (l1) {
            e1s = tks[lane + 64];
            e1i = tki[lane + 64];
        }
        for (uint64_t m = cm; m != 0; m &= m - 1) {
            const int bl = __ffsll(static_cast<unsigned long long>(m)) - 1;
            const float cs = lane_f(sc, bl);
            const int32_t ci = __builtin_amdgcn_readlane(id, bl);
            const int cnt = __popcll(ballot(l0 && better(e0s, e0i, cs, ci))) +
                            __popcll(ballot(l1 && better(e1s, e1i, cs, ci)));
            const bool dp = ballot((l0 && e0i == ci) || (l1 && e1i == ci)) != 0;
            if (lane == bl) {
                rt = cnt;
                dup = dp;
            }
        }
    } else if (cand) {
        int hi = tkn;
        while (rt < hi) {
            const int mid = (rt + hi) >> 1;
            if (better(tks[mid], tki[mid], sc, id))
                rt = mid + 1;
            else
                hi = mid;
        }
        dup = rt < tkn && tki[rt] == id;
    }
    const bool v = cand && !dup;
    const uint64_t vm = ballot(v);
    if (vm == 0) return;
    int rc = 0;
    for (uint64_t m = vm; m != 0; m &= m - 1) {
        const int bl = __ffsll(static_cast<unsigned long long>(m)) - 1;
        rc += better(lane_f(sc, bl), __builtin_amdgcn_readlane(id, bl), sc, id) ? 1 : 0;
    }
    // top-k entries move up past the candidates better than them, highest
    // tile first (a tile is read whole before it is written)
    for (int t0 = ((tkn - 1) >> 6) << 6; t0 >= 0 && tkn > 0; t0 -= 64) {
        const int t = t0 + lane;
        const bool live = t < tkn;
        float es;
        int32_t ei;
        if (small) {
            es = t0 == 0 ? e0s : e1s;
            ei = t0 == 0 ? e0i : e1i;
        } else {
            es = live ? tks[t] : 0.0F;
            ei = live ? tki[t] : 0;
        }
        int up = 0;
        for (uint64_t m = vm; m != 0; m &= m - 1) {
            const int bl = __ffsll(static_cast<unsigned long long>(m)) - 1;
            up += better(lane_f(sc, bl), __builtin_amdgcn_readlane(id, bl), es, ei) ? 1 : 0;
        }
        __builtin_amdgcn_wave_barrier();
        if (live && up > 0 && t + up < k) {
            tks[t + up] = es;
            tki[t + up] = ei;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (v && rt + rc < k) {
        tks[rt + rc] = sc;
        tki[rt + rc] = id;
    }
    __builtin_amdgcn_wave_barrier();
    tkn = min(k, tkn + __popcll(vm));
}

// LDS tables of one query (all pointers into the workgroup's LDS).
struct QueryTables {
    uint32_t* qbits;  // presence bitmap over the vocabulary (nw words)
    uint16_t* qpref;  // set bits in the words before each word
    uint16_t* eterm;  // per query entry: term id
    uint16_t* erank;  //   its rank among the distinct terms
    float* evalf;     //   its uint8 code (SQ) or value
    int32_t* rlast;   // per distinct term: its last entry
    float* rval;      //   that entry's code or value (the reference's dense vector)
    uint32_t* rsum;   //   sum of its entries' codes (SQ)
    uint8_t* qcode;   // DENSE: vocabulary-sized code table (last entry wins)
    int32_t* wsum;    // one int per wave
    int32_t* err;     // zero before the call
    int32_t* nu;      // distinct terms, written here
};

// Loads query entries [s0, s0 + qn) into the tables. Returns false (uniformly,
// after a barrier) when a term id is >= dim; the tables are then partial.
// Codes are ScalarQuantizer::encode_8bit (SQ) or encode_16bit (W16) of the
// values (utils/scalar_quantizer.h:97-113), held as exact floats. No barrier
// follows the last writes.
template <bool SQ, bool DENSE, int BLOCK, bool W16 = false>
__device__ __forceinline__ bool load_query_lds(const SearchArgs& a, int64_t s0, int qn, int nw,
                                               const QueryTables& T) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int w = tid; w < nw; w += BLOCK) T.qbits[w] = 0;
    for (int r = tid; r < qn; r += BLOCK) {
        T.rlast[r] = -1;
        T.rsum[r] = 0;
    }
    __syncthreads();
    for (int i = tid; i < qn; i += BLOCK) {
        const uint32_t t = a.q_indices[s0 + i];
        const float v = a.q_values[s0 + i];
        if (t >= static_cast<uint32_t>(a.ix.dim)) {
            *T.err = 1;
            continue;
        }
        T.eterm[i] = static_cast<uint16_t>(t);
        if constexpr (SQ || W16) {
            constexpr float kMax = W16 ? 65535.0F : 255.0F;
            float sc = (v - a.q_vmin) * a.q_scale;
            sc = (sc < kMax) ? sc : kMax;
            sc = (0.0F < sc) ? sc : 0.0F;
            T.evalf[i] = static_cast<float>(static_cast<uint32_t>(roundf(sc)));
        } else {
            T.evalf[i] = v;
        }
        atomicOr(&T.qbits[t >> 5], 1u << (t & 31));
    }
    __syncthreads();
    if (*T.err != 0) return false;
    // exclusive popcount prefix over the bitmap words
    {
        const int per = (nw + BLOCK - 1) / BLOCK;
        const int w0 = tid * per;
        int local = 0;
        for (int w = w0; w < w0 + per && w < nw; ++w) local += __popc(T.qbits[w]);
        const int incl = wave_incl_scan(local, lane);
        if (lane == 63) T.wsum[wave] = incl;
        __syncthreads();
        int base = 0;
        for (int w = 0; w < wave; ++w) base += T.wsum[w];
        int run = base + incl - local;
        for (int w = w0; w < w0 + per && w < nw; ++w) {
            T.qpref[w] = static_cast<uint16_t>(run);
            run += __popc(T.qbits[w]);
        }
        if (tid == BLOCK - 1) *T.nu = base + incl;
    }
    __syncthreads();
    const QueryLds QL{T.qbits, T.qpref, T.rval};
    for (int i = tid; i < qn; i += BLOCK) {
        bool p;
        const int r = QL.rank_of(T.eterm[i], p);
        T.erank[i] = static_cast<uint16_t>(r);
        atomicMax(&T.rlast[r], i);
        if constexpr (SQ) atomicAdd(&T.rsum[r], static_cast<uint32_t>(T.evalf[i]));
    }
    __syncthreads();
    const int nu = *T.nu;
    for (int r = tid; r < nu; r += BLOCK) T.rval[r] = T.evalf[T.rlast[r]];
    if constexpr (DENSE) {
        uint32_t* q32 = reinterpret_cast<uint32_t*>(T.qcode);
        for (int w = tid; w < (a.ix.dim + 3) / 4; w += BLOCK) q32[w] = 0;
        __syncthreads();
        // get_dense_vector: the last entry of a repeated term wins
        for (int i = tid; i < qn; i += BLOCK)
            if (T.rlast[T.erank[i]] == i) T.qcode[T.eterm[i]] = static_cast<uint8_t>(T.evalf[i]);
    }
    return true;
}

// The summary columns of list `term` (summary terms [tb, te)) that the query's
// terms select: rcol[rank of a distinct query term] = its column, left as the
// caller preset it (-1) for terms the list's summaries lack. With per-list
// summary-term bitmaps each query entry reads one bitmap word and one rank
// prefix (a repeated term writes the same column); without, the list's term
// ids are scanned against the query bitmap.
template <int BLOCK>
__device__ __forceinline__ void match_summary_columns(const DevIndexView& ix, uint32_t term, int32_t tb, int32_t te,
                                                      int qn, const uint16_t* eterm, const uint16_t* erank,
                                                      const QueryLds& QL, int32_t* rcol) {
    const int tid = threadIdx.x;
    if (ix.sum_bits != nullptr) {
        const uint32_t* lbits = ix.sum_bits + static_cast<int64_t>(term) * ix.sum_nw;
        const uint16_t* lrank = ix.sum_rank + static_cast<int64_t>(term) * ix.sum_nw;
        for (int i = tid; i < qn; i += BLOCK) {
            const uint32_t t = eterm[i];
            const uint32_t w = lbits[t >> 5], bit = 1u << (t & 31);
            const int32_t below = lrank[t >> 5];
            if (w & bit) rcol[erank[i]] = tb + below + __popc(w & (bit - 1));
        }
    } else {
        constexpr int kU = 4;  // independent term loads in flight per lane
        for (int i0 = tb + tid; i0 < te; i0 += BLOCK * kU) {
            uint32_t t[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int i = i0 + u * BLOCK;
                t[u] = i < te ? ix.sum_term_ids[i] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (t[u] == 0xFFFFFFFFu) continue;
                bool p;
                const int r = QL.rank_of(t[u], p);
                if (p) rcol[r] = i0 + u * BLOCK;
            }
        }
    }
}

// Integer summary scores of one list (uint8 codes; InvertedListClusters::
// score_summaries_transposed, cluster/inverted_list_clusters.cpp:295-331):
// sacc[cluster] += rsum[r] * code over every entry of the found columns
// (flist[0, nf) holds their distinct-term ranks, rcol their columns). The
// entries are flattened: one pass loads every column's range and query-code
// sum, a prefix over their lengths numbers the entries, then all lanes take
// entries with kU loads in flight, each finding its column by binary search.
// Integer atomics make the per-cluster sums exact in any order. `uni` holds
// qcap * 12 + 4 bytes of scratch (qcap >= the query's distinct terms). Ends
// with a barrier.
template <int BLOCK>
__device__ __forceinline__ void accumulate_summaries_sq(const DevIndexView& ix, int32_t tb, int32_t te, int nf,
                                                        const int32_t* flist, const int32_t* rcol,
                                                        const uint32_t* rsum, unsigned char* uni, uint32_t* sacc,
                                                        int qcap = kQMax) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint8_t* sv = static_cast<const uint8_t*>(ix.sum_value);
    int32_t* fe0 = reinterpret_cast<int32_t*>(uni);  // column start, relative to the list's first entry
    int32_t* fpre = fe0 + qcap;                       // [nf + 1] exclusive prefix of column lengths
    uint32_t* fqs = reinterpret_cast<uint32_t*>(fpre + qcap + 1);
    const int64_t lbase = te > tb ? ix.sum_term_ptr[tb] : 0;
    for (int f = tid; f < nf; f += BLOCK) {
        const int r = flist[f];
        const int col = rcol[r];
        const int64_t e0 = ix.sum_term_ptr[col], e1 = ix.sum_term_ptr[col + 1];
        fe0[f] = static_cast<int32_t>(e0 - lbase);
        fpre[f + 1] = static_cast<int32_t>(e1 - e0);
        fqs[f] = rsum[r];
    }
    __syncthreads();
    if (wave == 0) {
        int run = 0;
        for (int b = 0; b < nf; b += 64) {
            const int f = b + lane;
            const int len = f < nf ? fpre[f + 1] : 0;
            const int incl = wave_incl_scan(len, lane);
            if (f < nf) fpre[f + 1] = run + incl;
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) fpre[0] = 0;
    }
    __syncthreads();
    const int total_e = fpre[nf];
    constexpr int kU = 4;
    for (int x0 = tid; x0 < total_e; x0 += BLOCK * kU) {
        int fx[kU];
        int64_t ex[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int x = x0 + u * BLOCK;
            fx[u] = -1;
            ex[u] = lbase;
            if (x < total_e) {
                int lo = 0, hi = nf - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (fpre[mid] <= x)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                fx[u] = lo;
                ex[u] = lbase + fe0[lo] + (x - fpre[lo]);
            }
        }
        uint32_t cl[kU], vv[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            cl[u] = ix.sum_cluster[ex[u]];
            vv[u] = sv[ex[u]];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (fx[u] >= 0) atomicAdd(&sacc[cl[u]], fqs[fx[u]] * vv[u]);
    }
    __syncthreads();
}

}  // namespace nsparse_amd
