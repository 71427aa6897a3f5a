// GPU side of add() and build(): scalar quantization and the SEISMIC index
// build (build_inverted_lists_clusters, seismic_common.h:146-184) as HIP
// kernels on one device.
//
// Phases (all in HBM; only per-list sizes come back to the host):
//  1. postings: stable radix sort of the forward index by term (doc order kept
//     inside a term) -- ArrayInvertedLists::build_inverted_lists;
//  2. lambda pruning: a second stable sort by (term, value desc) gives each
//     long list's kept docs in value order, ties by doc id --
//     InvertedList::prune_and_keep_doc_ids;
//  3. centroid draws: std::mt19937(seed + list) with libstdc++'s
//     uniform_int_distribution (Lemire) and the swap-on-a-copy of
//     random_select_initial_centroids, one lane per list;
//  4. assignment: one wavefront per document walks the document's terms in CSR
//     order against a term-major index of the list's centroids, adding into an
//     LDS similarity row (int for uint8 codes, float in term order for FP32) and
//     taking the strict-> argmax -- map_docs_to_clusters;
//  5. clusters: centroid first, then assigned docs in pruned order;
//  6. summaries: per-cluster max-pool and alpha-mass cut (summarize_*), uint8
//     via an LDS byte-max table and a 256-bin value histogram, FP32 with the
//     reference's sequential float sums;
//  7. transpose: stable sort of the summary entries by (list, term) gives the
//     term-major CSC of every list -- InvertedListClusters::build_transpose.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <random>
#include <type_traits>
#include <stdexcept>
#include <vector>

#include "device_mem.h"
#include "kernels.h"
#include "stl_order.h"

namespace nsparse_amd {

namespace {

constexpr int kB = 256;

unsigned grid_for(size_t n, size_t per = kB) {
    return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>((n + per - 1) / per, 1u << 20)));
}

#define GRID_STRIDE(i, n) \
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < (n); \
         i += static_cast<size_t>(gridDim.x) * blockDim.x)

template <class K, class V>
void radix_sort_pairs(const K* kin, K* kout, const V* vin, V* vout, size_t n, int end_bit) {
    if (n == 0) return;
    size_t tmp = 0;
    hip_check(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, n, 0, end_bit), "radix sort (size)");
    DeviceBuffer t(tmp);
    hip_check(rocprim::radix_sort_pairs(t.get(), tmp, kin, kout, vin, vout, n, 0, end_bit), "radix sort");
}

// out[i] = sum(in[0..i)), out[n] = total; `in` and `out` are int64.
void exclusive_scan_i64(const int64_t* in, int64_t* out, size_t n) {
    if (n == 0) {
        hip_check(hipMemset(out, 0, 8), "hipMemset");
        return;
    }
    size_t tmp = 0;
    hip_check(rocprim::inclusive_scan(nullptr, tmp, in, out + 1, n, rocprim::plus<int64_t>()), "scan (size)");
    DeviceBuffer t(tmp);
    hip_check(rocprim::inclusive_scan(t.get(), tmp, in, out + 1, n, rocprim::plus<int64_t>()), "scan");
    hip_check(hipMemset(out, 0, 8), "hipMemset");
}

void sync(const char* what) {
    hip_check(hipGetLastError(), what);
    hip_check(hipDeviceSynchronize(), what);
}

// ------------------------------------------------------------------ kernels

// ScalarQuantizer::encode_8bit / encode_16bit (utils/scalar_quantizer.h:97-113):
// the scale is computed on the host as `kMax / (vmax - vmin)` (kMax 255 or
// 65535), exactly as the reference evaluates it; the device applies
// (v - vmin) * scale, clamps with std::min/std::max semantics and rounds half
// away from zero (lround).
template <class T>
__global__ void sq_encode_kernel(const float* __restrict__ v, T* __restrict__ out, size_t n, float vmin, float scale) {
    constexpr float kMax = sizeof(T) == 1 ? 255.0F : 65535.0F;
    GRID_STRIDE(i, n) {
        float s = (v[i] - vmin) * scale;
        s = (s < kMax) ? s : kMax;
        s = (0.0F < s) ? s : 0.0F;
        out[i] = static_cast<T>(static_cast<uint32_t>(roundf(s)));
    }
}

__global__ void iota_kernel(int32_t* v, size_t n) {
    GRID_STRIDE(i, n) v[i] = static_cast<int32_t>(i);
}

// Postings per term from the term-sorted keys: each run adds its end and
// subtracts its start (hist is zero on entry).
__global__ void term_runs_kernel(const uint16_t* __restrict__ keys, size_t n, int32_t* __restrict__ hist) {
    GRID_STRIDE(i, n) {
        const uint16_t t = keys[i];
        if (i == 0 || keys[i - 1] != t) atomicSub(&hist[t], static_cast<int32_t>(i));
        if (i + 1 == n || keys[i + 1] != t) atomicAdd(&hist[t], static_cast<int32_t>(i + 1));
    }
}

__global__ void doc_of_kernel(const int32_t* __restrict__ indptr, int64_t n_docs, int32_t* __restrict__ doc_of) {
    GRID_STRIDE(d, static_cast<size_t>(n_docs)) {
        for (int32_t p = indptr[d]; p < indptr[d + 1]; ++p) doc_of[p] = static_cast<int32_t>(d);
    }
}

// Pruning key: term in the high bits, value descending below it.
__global__ void prune_key_u8_kernel(const uint16_t* __restrict__ idx, const uint8_t* __restrict__ v, size_t n,
                                    uint32_t* __restrict__ key) {
    GRID_STRIDE(i, n) key[i] = (static_cast<uint32_t>(idx[i]) << 8) | (255u - v[i]);
}

__global__ void prune_key_u16_kernel(const uint16_t* __restrict__ idx, const uint16_t* __restrict__ v, size_t n,
                                     uint32_t* __restrict__ key) {
    GRID_STRIDE(i, n) key[i] = (static_cast<uint32_t>(idx[i]) << 16) | (65535u - v[i]);
}

__device__ __forceinline__ uint32_t float_desc_bits(float f) {
    if (f == 0.0F) f = 0.0F;  // -0.0 and +0.0 compare equal in the reference
    const uint32_t b = __float_as_uint(f);
    const uint32_t asc = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ~asc;
}

__global__ void prune_key_f32_kernel(const uint16_t* __restrict__ idx, const float* __restrict__ v, size_t n,
                                     uint64_t* __restrict__ key) {
    GRID_STRIDE(i, n) key[i] = (static_cast<uint64_t>(idx[i]) << 32) | float_desc_bits(v[i]);
}

// kept docs of every list: doc-ordered postings when the whole list is kept,
// value-ordered prefix otherwise.
__global__ void kept_docs_kernel(const int64_t* __restrict__ list_start, const int64_t* __restrict__ kept_base,
                                 const int32_t* __restrict__ spos, const int32_t* __restrict__ spos2,
                                 const int32_t* __restrict__ doc_of, int32_t* __restrict__ kept, int n_lists) {
    for (int l = blockIdx.x; l < n_lists; l += gridDim.x) {
        const int64_t s = list_start[l], len = list_start[l + 1] - s;
        const int64_t kb = kept_base[l], kl = kept_base[l + 1] - kb;
        const int32_t* src = (kl < len) ? spos2 : spos;
        for (int64_t i = threadIdx.x; i < kl; i += blockDim.x) kept[kb + i] = doc_of[src[s + i]];
    }
}

// std::mt19937 (standard-specified) and libstdc++'s uniform_int_distribution
// for a 32-bit engine (Lemire's nearly-divisionless method with a 64-bit
// product, bits/uniform_int_dist.h _S_nd).
struct Mt19937 {
    uint32_t* mt;
    int idx;
    __device__ void seed(uint32_t s) {
        mt[0] = s;
        for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
        idx = 624;
    }
    __device__ uint32_t next() {
        if (idx >= 624) {
            for (int i = 0; i < 624; ++i) {
                const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            idx = 0;
        }
        uint32_t y = mt[idx++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    // uniform_int_distribution<int>(0, range - 1)
    __device__ uint32_t below(uint32_t range) {
        uint64_t product = static_cast<uint64_t>(next()) * range;
        uint32_t low = static_cast<uint32_t>(product);
        if (low < range) {
            const uint32_t threshold = (0u - range) % range;
            while (low < threshold) {
                product = static_cast<uint64_t>(next()) * range;
                low = static_cast<uint32_t>(product);
            }
        }
        return static_cast<uint32_t>(product >> 32);
    }
};

// random_select_initial_centroids (random_kmeans.cpp:28-51) on a copy of the
// kept docs; writes centroid doc ids in draw order.
__global__ void centroid_kernel(const int64_t* __restrict__ kept_base, const int32_t* __restrict__ kept,
                                int32_t* __restrict__ work, const int32_t* __restrict__ lcb,
                                int32_t* __restrict__ centroid_doc, uint32_t base_seed, int n_lists) {
    __shared__ uint32_t state[624];
    for (int l = blockIdx.x; l < n_lists; l += gridDim.x) {
        const int64_t kb = kept_base[l];
        const int64_t n = kept_base[l + 1] - kb;
        const int nc = lcb[l + 1] - lcb[l];
        if (n == 0 || nc == 0) continue;
        for (int64_t i = threadIdx.x; i < n; i += blockDim.x) work[kb + i] = kept[kb + i];
        __syncthreads();
        if (threadIdx.x == 0) {
            Mt19937 g{state, 0};
            g.seed(base_seed + static_cast<uint32_t>(l));
            for (int i = 0; i < nc; ++i) {
                const uint32_t r = g.below(static_cast<uint32_t>(n - i));
                const int32_t d = work[kb + r];
                centroid_doc[lcb[l] + i] = d;
                work[kb + r] = work[kb + n - 1 - i];
                work[kb + n - 1 - i] = d;
            }
        }
        __syncthreads();
    }
}

__global__ void cluster_list_kernel(const int32_t* __restrict__ lcb, int32_t* __restrict__ cluster_list, int n_lists) {
    for (int l = blockIdx.x; l < n_lists; l += gridDim.x)
        for (int c = lcb[l] + threadIdx.x; c < lcb[l + 1]; c += blockDim.x) cluster_list[c] = l;
}

// Row length of each cluster's centroid; zero for clusters of lists whose
// every kept doc is a centroid (nothing is assigned there, so their centroids
// need no term-major index).
__global__ void row_len_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ docs,
                               const int32_t* __restrict__ cluster_list, const int32_t* __restrict__ lcb,
                               const int64_t* __restrict__ kept_base, size_t n, int64_t* __restrict__ out) {
    GRID_STRIDE(i, n) {
        const int l = cluster_list[i];
        const bool needed = (kept_base[l + 1] - kept_base[l]) > (lcb[l + 1] - lcb[l]);
        out[i] = needed ? indptr[docs[i] + 1] - indptr[docs[i]] : 0;
    }
}

// One posting per (centroid term): key (list, term), value = entry id. Covers
// clusters [g0, g0 + n_clusters); entry ids are relative to e_base, the first
// entry of cluster g0, so one batch of lists fits 32-bit sort values. A group
// of 16 lanes writes one centroid's row, so the stores are contiguous.
template <class T>
__global__ __launch_bounds__(kB) void centroid_entries_kernel(
    const int32_t* __restrict__ indptr, const uint16_t* __restrict__ indices, const T* __restrict__ values,
    const int32_t* __restrict__ centroid_doc, const int32_t* __restrict__ cluster_list, const int32_t* __restrict__ lcb,
    const int64_t* __restrict__ cent_off, size_t g0, size_t n_clusters, int64_t e_base, uint32_t* __restrict__ key,
    int32_t* __restrict__ val, uint16_t* __restrict__ ent_c, T* __restrict__ ent_w) {
    const int grp = threadIdx.x >> 4, gl = threadIdx.x & 15;
    for (size_t i = static_cast<size_t>(blockIdx.x) * (kB / 16) + grp; i < n_clusters;
         i += static_cast<size_t>(gridDim.x) * (kB / 16)) {
        const size_t g = g0 + i;
        const int64_t e0 = cent_off[g] - e_base;
        if (cent_off[g + 1] - e_base == e0) continue;  // list without docs to assign
        const int32_t d = centroid_doc[g];
        const int l = cluster_list[g];
        const uint16_t c = static_cast<uint16_t>(g - lcb[l]);
        const int32_t r0 = indptr[d], len = indptr[d + 1] - r0;
        for (int j = gl; j < len; j += 16) {
            const int64_t e = e0 + j;
            key[e] = (static_cast<uint32_t>(l) << 16) | indices[r0 + j];
            val[e] = static_cast<int32_t>(e);
            ent_c[e] = c;
            ent_w[e] = values[r0 + j];
        }
    }
}

template <class T>
__global__ void gather_csc_kernel(const int32_t* __restrict__ perm, const uint16_t* __restrict__ ent_c,
                                  const T* __restrict__ ent_w, size_t n, uint16_t* __restrict__ c_out,
                                  T* __restrict__ w_out) {
    GRID_STRIDE(i, n) {
        c_out[i] = ent_c[perm[i]];
        w_out[i] = ent_w[perm[i]];
    }
}

__global__ void fill_i32_kernel(int32_t* p, size_t n, int32_t v) {
    GRID_STRIDE(i, n) p[i] = v;
}

// Inclusive prefix sum over the 64 lanes with DPP moves only: Hillis-Steele
// inside each 16-lane row (row_shr 1, 2, 4, 8), then row_bcast:15 and
// row_bcast:31 carry the row totals upward.
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false));
    return v;
}

template <class Acc>
__device__ __forceinline__ void argmax_merge(Acc& v, int& i, Acc ov, int oi) {
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

// map_docs_to_clusters_typed (cluster/kmeans_utils.cpp:107-158). One
// workgroup per list of [l_begin, l_end) (grid-stride), one wavefront per
// document. `ckey`/`cc`/`cw` hold the batch's centroid CSC, whose entry ids are
// global ids minus e_base. `tmap` is this workgroup's dense term -> [begin, end)
// table of the list's centroid CSC (batch-relative entry ids).
template <class T, class Acc>
__global__ __launch_bounds__(kB) void assign_kernel(
    const int32_t* __restrict__ indptr, const uint16_t* __restrict__ indices, const T* __restrict__ values,
    const int64_t* __restrict__ kept_base, const int32_t* __restrict__ kept, const int32_t* __restrict__ lcb,
    const int32_t* __restrict__ centroid_doc, const int64_t* __restrict__ cent_off, int64_t e_base,
    const uint32_t* __restrict__ ckey, const uint16_t* __restrict__ cc, const T* __restrict__ cw,
    int2* __restrict__ tmap_all, int dim, int32_t* __restrict__ assign, int l_begin, int l_end, int mc,
    unsigned long long* __restrict__ pairs) {
    extern __shared__ __align__(16) unsigned char smem[];
    int mc2 = 1;  // the sort's power-of-two capacity
    while (mc2 < mc) mc2 <<= 1;
    int32_t* cent = reinterpret_cast<int32_t*>(smem);                      // sorted centroid ids [mc2]
    Acc* sims_all = reinterpret_cast<Acc*>(smem + ((mc2 * 4 + 15) & ~15));  // [4][mc]
    int2* tmap = tmap_all + static_cast<size_t>(blockIdx.x) * dim;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Acc* sims = sims_all + wave * mc;
    uint64_t n_pairs = 0;  // (doc term, centroid posting) pairs this wave scored
    for (int l = l_begin + blockIdx.x; l < l_end; l += gridDim.x) {
        const int64_t kb = kept_base[l], n = kept_base[l + 1] - kb;
        const int c0 = lcb[l], nc = lcb[l + 1] - c0;
        if (n == 0 || nc == 0) continue;
        if (n == nc) {  // every kept doc is a centroid: nothing to assign
            for (int64_t j = tid; j < n; j += kB) assign[kb + j] = -1;
            continue;
        }
        const int64_t e0 = cent_off[c0] - e_base, e1 = cent_off[c0 + nc] - e_base;
        for (int64_t e = e0 + tid; e < e1; e += kB) {
            const uint32_t k = ckey[e];
            const int t = static_cast<int>(k & 0xFFFFu);
            if (e == e0 || ckey[e - 1] != k) tmap[t].x = static_cast<int>(e);
            if (e == e1 - 1 || ckey[e + 1] != k) tmap[t].y = static_cast<int>(e + 1);
        }
        // centroid ids, sorted ascending (bitonic, padded to a power of two)
        int n2 = 1;
        while (n2 < nc) n2 <<= 1;
        for (int i = tid; i < n2; i += kB) cent[i] = i < nc ? centroid_doc[c0 + i] : INT32_MAX;
        __syncthreads();
        for (int size = 2; size <= n2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = tid; t < n2 / 2; t += kB) {
                    const int i = 2 * t - (t & (stride - 1)), j = i + stride;
                    const int32_t a = cent[i], b = cent[j];
                    if ((a > b) == ((i & size) == 0)) {
                        cent[i] = b;
                        cent[j] = a;
                    }
                }
                __syncthreads();
            }
        }
        for (int64_t j = wave; j < n; j += kB / 64) {
            const int32_t d = kept[kb + j];
            int lo = 0, hi = nc;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cent[mid] < d)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < nc && cent[lo] == d) {
                if (lane == 0) assign[kb + j] = -1;
                continue;
            }
            for (int c = lane; c < nc; c += 64) sims[c] = Acc(0);
            __builtin_amdgcn_wave_barrier();
            const int32_t r0 = indptr[d], r1 = indptr[d + 1];
            if constexpr (!std::is_same_v<T, float>) {
                // Integer sims are exact in any order: the doc's (term, centroid
                // posting) pairs are numbered by a prefix over the terms' posting
                // counts and dealt to the lanes, kW windows of 64 with their loads
                // in flight together, and added with LDS atomics. A lane finds
                // its term as the last lane whose prefix is <= its item.
                constexpr int kW = 16;
                for (int b = r0; b < r1; b += 64) {
                    const int p = b + lane;
                    int cs = 0, cnt = 0;
                    Acc dv = Acc(0);
                    if (p < r1) {
                        const int2 col = tmap[indices[p]];
                        if (col.x >= 0) {
                            cs = col.x;
                            cnt = col.y - col.x;
                        }
                        dv = static_cast<Acc>(values[p]);
                    }
                    const int incl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(cnt)));
                    const int excl = incl - cnt;
                    const int total = __builtin_amdgcn_readlane(incl, 63);
                    n_pairs += static_cast<uint64_t>(total);
                    for (int i0 = 0; i0 < total; i0 += 64 * kW) {
                        int ev[kW];
                        Acc wv[kW];
                        bool ok[kW];
#pragma unroll
                        for (int u = 0; u < kW; ++u) {
                            const int w0 = i0 + 64 * u, i = w0 + lane;
                            ok[u] = i < total;
                            ev[u] = 0;
                            wv[u] = Acc(0);
                            if (w0 >= total) continue;  // wave-uniform
                            // lanes with excl <= w0 form a prefix whose last lane has postings
                            int owner = __popcll(__ballot(excl <= w0)) - 1;
                            const uint64_t bm = __ballot(excl > w0 && excl <= w0 + 63 && cnt > 0);
                            for (uint64_t m = bm; m != 0; m &= m - 1) {
                                const int l = __ffsll(static_cast<unsigned long long>(m)) - 1;
                                if (i >= __builtin_amdgcn_readlane(excl, l)) owner = l;
                            }
                            const int ocs = __shfl(cs, owner, 64), oex = __shfl(excl, owner, 64);
                            wv[u] = __shfl(dv, owner, 64);
                            ev[u] = ocs + (i - oex);
                        }
                        uint32_t cl[kW], cwv[kW];
#pragma unroll
                        for (int u = 0; u < kW; ++u) {
                            cl[u] = ok[u] ? cc[ev[u]] : 0u;
                            cwv[u] = ok[u] ? static_cast<uint32_t>(cw[ev[u]]) : 0u;
                        }
#pragma unroll
                        for (int u = 0; u < kW; ++u)
                            if (ok[u]) atomicAdd(&sims[cl[u]], wv[u] * static_cast<Acc>(cwv[u]));
                    }
                }
                __builtin_amdgcn_wave_barrier();
            } else {
                // FP32 sims add in doc-term order, as the reference does: the
                // terms one after another, lanes over a term's (distinct) clusters
                for (int b = r0; b < r1; b += 64) {
                    const int p = b + lane;
                    int2 col = make_int2(0, 0);
                    Acc dv = Acc(0);
                    if (p < r1) {
                        col = tmap[indices[p]];
                        dv = static_cast<Acc>(values[p]);
                    }
                    const int m = min(64, r1 - b);
                    for (int j2 = 0; j2 < m; ++j2) {
                        const int cs = __shfl(col.x, j2, 64), ce = __shfl(col.y, j2, 64);
                        const Acc v = __shfl(dv, j2, 64);
                        if (cs < 0) continue;
                        n_pairs += static_cast<uint64_t>(ce - cs);
                        for (int e = cs + lane; e < ce; e += 64) sims[cc[e]] += v * static_cast<Acc>(cw[e]);
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
            Acc best = sims[0];
            int bi = 0;
            for (int c = lane; c < nc; c += 64) argmax_merge(best, bi, sims[c], c);
            for (int off = 32; off > 0; off >>= 1) {
                const Acc ov = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                argmax_merge(best, bi, ov, oi);
            }
            if (lane == 0) assign[kb + j] = bi;
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        for (int64_t e = e0 + tid; e < e1; e += kB) tmap[ckey[e] & 0xFFFFu] = make_int2(-1, -1);
        __syncthreads();
    }
    if (lane == 0 && n_pairs != 0) atomicAdd(pairs, static_cast<unsigned long long>(n_pairs));
}

// ---- uint8 assignment with a dense part on the matrix cores
//
// map_docs_to_clusters_typed (cluster/kmeans_utils.cpp:107-158) for uint8
// codes. A list's (doc term, centroid posting) pairs concentrate on the few
// terms that most of its centroids share: at 8.8M synthetic docs the 256 terms
// with the most centroid postings carry about 96% of a list's pairs. For those
// "hot" terms sims[d][c] = sum_t doc[d][t] * cent[c][t] is a dense product of a
// (docs x hot terms) tile and a (hot terms x centroids) matrix, computed with
// v_mfma_f32_16x16x32_bf16: codes (< 256) are exact in bfloat16, every product
// (< 2^16) and every partial sum over at most 256 terms (< 2^24) is an exact
// float, so the dense part is the exact integer in any order. The remaining
// "cold" pairs are added with integer LDS atomics. The sum of both is the
// reference's integer similarity; the argmax keeps the strict > (lowest
// cluster on ties, cluster 0 when all are zero).

constexpr int kHot = 256;        // hot term slots (the dense product's K)
constexpr int kTileDocs = 16;    // docs per dense tile (the MFMA's M)
constexpr int kHotHist = 1024;   // centroid-posting count histogram bins (counts clamp)
constexpr int kHB = 512;         // hot assignment workgroup
constexpr int kColdCap = 2048;   // cold (doc term) entries of a tile staged at once
constexpr int kHotMaxClusters = 831;  // largest list the LDS plan takes

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// Dynamic LDS of the hot assignment for lists of up to mc clusters: sorted
// centroid ids, the hot tile, the tile's sims (or the count histogram), the
// cold list (term range start, prefix of posting counts, row | code << 8) and
// per-tile doc tables.
struct HotLds {
    int o_cent, o_tile, o_sims, o_ccs, o_cpre, o_crd, o_tdoc, o_tstart, o_tpre, o_misc, total;
};
__host__ __device__ inline HotLds hot_lds(int mc) {
    int mc2 = 1;
    while (mc2 < mc) mc2 <<= 1;
    HotLds L{};
    int o = 0;
    auto take = [&](int bytes) {
        const int at = o;
        o += (bytes + 15) & ~15;
        return at;
    };
    L.o_cent = take(mc2 * 4);
    L.o_tile = take(kTileDocs * kHot * 2);
    L.o_sims = take((kTileDocs * mc > kHotHist ? kTileDocs * mc : kHotHist) * 4);
    L.o_ccs = take(kColdCap * 4);
    L.o_cpre = take((kColdCap + 1) * 4);
    L.o_crd = take(kColdCap * 4);
    L.o_tdoc = take(kTileDocs * 4);
    L.o_tstart = take(kTileDocs * 4);
    L.o_tpre = take((kTileDocs + 1) * 4);
    L.o_misc = take(16 * 4);
    L.total = o;
    return L;
}

// One 512-lane workgroup per list of [l_begin, l_end) (grid-stride). `tmap`
// is this workgroup's term -> CSC range table (as in assign_kernel); during a
// list a hot term's entry holds (-2 - slot, its centroid postings). `bmat` is
// this workgroup's (centroids x kHot) bfloat16 matrix of the hot terms' codes,
// rows padded to a multiple of 16. Lists have at most mc <= kHotMaxClusters
// clusters; dynamic LDS hot_lds(mc).total bytes.
//
// Per tile of 16 docs: (1) every lane takes (doc, term) entries of the tile's
// rows, four at a time with their loads in flight; a hot term's code goes to
// the tile, a cold term with centroid postings joins the cold list; (2) the
// cold list's postings are numbered by a prefix over their counts and dealt to
// every lane, eight pairs at a time, each finding its entry by binary search,
// and added into the tile's sims with LDS atomics (rounds of at most kColdCap
// entries); (3) the dense part; (4) the argmax.
__global__ __launch_bounds__(kHB, 4) void assign_hot_kernel(
    const int32_t* __restrict__ indptr, const uint16_t* __restrict__ indices, const uint8_t* __restrict__ values,
    const int64_t* __restrict__ kept_base, const int32_t* __restrict__ kept, const int32_t* __restrict__ lcb,
    const int32_t* __restrict__ centroid_doc, const int64_t* __restrict__ cent_off, int64_t e_base,
    const uint32_t* __restrict__ ckey, const uint16_t* __restrict__ cc, const uint8_t* __restrict__ cw,
    int2* __restrict__ tmap_all, int dim, __bf16* __restrict__ bmat_all, int32_t* __restrict__ assign, int l_begin,
    int l_end, int mc, unsigned long long* __restrict__ pairs) {
    extern __shared__ __align__(16) unsigned char smem[];
    const HotLds LP = hot_lds(mc);
    const int mcp = (mc + 15) & ~15;
    int32_t* cent = reinterpret_cast<int32_t*>(smem + LP.o_cent);     // sorted centroid ids
    __bf16* atile = reinterpret_cast<__bf16*>(smem + LP.o_tile);        // [kTileDocs][kHot]
    int32_t* sims = reinterpret_cast<int32_t*>(smem + LP.o_sims);      // [kTileDocs][nc]
    int32_t* hist = sims;                                              // [kHotHist], list set-up only
    int32_t* ccs = reinterpret_cast<int32_t*>(smem + LP.o_ccs);        // cold: CSC range start
    int32_t* cpre = reinterpret_cast<int32_t*>(smem + LP.o_cpre);      // cold: posting counts, then their prefix
    int32_t* crd = reinterpret_cast<int32_t*>(smem + LP.o_crd);        // cold: row | code << 8
    int32_t* tdoc = reinterpret_cast<int32_t*>(smem + LP.o_tdoc);      // tile row -> doc (-1: none or a centroid)
    int32_t* tstart = reinterpret_cast<int32_t*>(smem + LP.o_tstart);  // tile row -> first element
    int32_t* tpre = reinterpret_cast<int32_t*>(smem + LP.o_tpre);      // tile rows' element prefix
    int32_t* misc = reinterpret_cast<int32_t*>(smem + LP.o_misc);      // [0] hot slots, [1] threshold, [2] cold entries
    int2* tmap = tmap_all + static_cast<size_t>(blockIdx.x) * dim;
    __bf16* bmat = bmat_all + static_cast<size_t>(blockIdx.x) * mcp * kHot;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int kWaves = kHB / 64;
    uint64_t n_pairs = 0;  // this lane's pairs, hot and cold
    for (int l = l_begin + blockIdx.x; l < l_end; l += gridDim.x) {
        const int64_t kb = kept_base[l], n = kept_base[l + 1] - kb;
        const int c0 = lcb[l], nc = lcb[l + 1] - c0;
        if (n == 0 || nc == 0) continue;
        if (n == nc) {  // every kept doc is a centroid: nothing to assign
            for (int64_t j = tid; j < n; j += kHB) assign[kb + j] = -1;
            continue;
        }
        const int64_t e0 = cent_off[c0] - e_base, e1 = cent_off[c0 + nc] - e_base;
        for (int i = tid; i < kHotHist; i += kHB) hist[i] = 0;
        if (tid == 0) misc[0] = 0;
        for (int64_t e = e0 + tid; e < e1; e += kHB) {
            const uint32_t k = ckey[e];
            const int t = static_cast<int>(k & 0xFFFFu);
            if (e == e0 || ckey[e - 1] != k) tmap[t].x = static_cast<int>(e);
            if (e == e1 - 1 || ckey[e + 1] != k) tmap[t].y = static_cast<int>(e + 1);
        }
        __syncthreads();
        // histogram of the terms' centroid-posting counts
        for (int64_t e = e0 + tid; e < e1; e += kHB) {
            const uint32_t k = ckey[e];
            if (e == e0 || ckey[e - 1] != k) {
                const int2 col = tmap[k & 0xFFFFu];
                atomicAdd(&hist[min(col.y - col.x, kHotHist - 1)], 1);
            }
        }
        __syncthreads();
        // threshold: the smallest count >= 2 whose terms (those with at least
        // that many postings) fit the kHot slots; lane i holds bins 16i..16i+15
        if (wave == 0) {
            int bins[16], sum = 0;
#pragma unroll
            for (int b = 0; b < 16; ++b) {
                bins[b] = hist[lane * 16 + b];
                sum += bins[b];
            }
            // suffix sums over the lanes: a prefix over the reversed order
            const int rev = __shfl(sum, 63 - lane, 64);
            const int rincl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(rev)));
            const int suf = __shfl(rincl, 63 - lane, 64);                    // terms in bins >= 16 lane
            const int above = lane == 63 ? 0 : __shfl(suf, lane + 1, 64);  // terms in bins >= 16 (lane + 1)
            // each lane's smallest bin b with S(b) <= kHot, if any of its bins fits
            int acc = above, tl = 16 * (lane + 1);
            for (int b = 15; b >= 0; --b) {
                if (acc + bins[b] > kHot) break;
                acc += bins[b];
                tl = 16 * lane + b;
            }
            // the lowest lane whose bins above its own fit holds the threshold
            const uint64_t ok = __ballot(above <= kHot);
            int theta = __shfl(tl, __ffsll(static_cast<unsigned long long>(ok)) - 1, 64);  // lane 63 always fits
            if (theta >= kHotHist) theta = INT32_MAX;  // even the top bin's terms overflow: no hot terms
            if (lane == 0) misc[1] = theta < 2 ? 2 : theta;
        }
        __syncthreads();
        const int theta = misc[1];
        // hot terms: a slot each, marked in tmap as (-2 - slot, postings)
        for (int64_t e = e0 + tid; e < e1; e += kHB) {
            const uint32_t k = ckey[e];
            if (e == e0 || ckey[e - 1] != k) {
                const int t = static_cast<int>(k & 0xFFFFu);
                const int2 col = tmap[t];
                const int cnt = col.y - col.x;
                if (cnt >= theta) {
                    const int slot = atomicAdd(&misc[0], 1);  // < kHot by the threshold's choice
                    if (slot < kHot) tmap[t] = make_int2(-2 - slot, cnt);
                }
            }
        }
        // zero the rows of bmat this list uses (16-byte stores), and sort the
        // centroid ids ascending (bitonic, padded to a power of two)
        const int ncp = (nc + 15) & ~15;
        {
            uint4* bz = reinterpret_cast<uint4*>(bmat);
            const int nz = ncp * kHot / 8;
            for (int i = tid; i < nz; i += kHB) bz[i] = make_uint4(0, 0, 0, 0);
        }
        int n2 = 1;
        while (n2 < nc) n2 <<= 1;
        for (int i = tid; i < n2; i += kHB) cent[i] = i < nc ? centroid_doc[c0 + i] : INT32_MAX;
        __syncthreads();
        for (int size = 2; size <= n2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = tid; t < n2 / 2; t += kHB) {
                    const int i = 2 * t - (t & (stride - 1)), j = i + stride;
                    const int32_t a = cent[i], b = cent[j];
                    if ((a > b) == ((i & size) == 0)) {
                        cent[i] = b;
                        cent[j] = a;
                    }
                }
                __syncthreads();
            }
        }
        // the hot terms' centroid codes into bmat[centroid][slot]
        for (int64_t e = e0 + tid; e < e1; e += kHB) {
            const int2 col = tmap[ckey[e] & 0xFFFFu];
            if (col.x <= -2) bmat[static_cast<size_t>(cc[e]) * kHot + (-2 - col.x)] = static_cast<__bf16>(cw[e]);
        }
        __syncthreads();
        const int nhot = misc[0] < kHot ? misc[0] : kHot;
        for (int64_t j0 = 0; j0 < n; j0 += kTileDocs) {
            const int nt = n - j0 < kTileDocs ? static_cast<int>(n - j0) : kTileDocs;
            // the tile's docs (centroids and rows past the list are empty) and
            // the prefix of their row lengths
            if (wave == 0) {
                int d = -1, r0 = 0, len = 0;
                if (lane < nt) {
                    d = kept[kb + j0 + lane];
                    int lo = 0, hi = nc;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (cent[mid] < d)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    if (lo < nc && cent[lo] == d) d = -1;
                    if (d >= 0) {
                        r0 = indptr[d];
                        len = indptr[d + 1] - r0;
                    }
                }
                const int incl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(len)));
                if (lane < kTileDocs) {
                    tdoc[lane] = d;
                    tstart[lane] = r0;
                    tpre[lane] = incl - len;
                }
                if (lane == kTileDocs - 1) tpre[kTileDocs] = incl;
            }
            {
                uint4* az = reinterpret_cast<uint4*>(atile);
                for (int i = tid; i < kTileDocs * kHot / 8; i += kHB) az[i] = make_uint4(0, 0, 0, 0);
                for (int i = tid; i < kTileDocs * nc; i += kHB) sims[i] = 0;
            }
            __syncthreads();
            const int E = tpre[kTileDocs];
            // rounds of at most kColdCap entries, so the cold list never overflows
            for (int s0 = 0; s0 < E; s0 += kColdCap) {
                const int s1 = min(E, s0 + kColdCap);
                if (tid == 0) misc[2] = 0;
                __syncthreads();
                // (1) entries: hot codes into the tile, cold terms into the list
                constexpr int kA = 4;
                for (int f0 = s0; f0 < s1; f0 += kHB * kA) {
                    int rr[kA], term[kA], val[kA];
#pragma unroll
                    for (int u = 0; u < kA; ++u) {
                        const int f = f0 + u * kHB + tid;
                        rr[u] = -1;
                        term[u] = 0;
                        val[u] = 0;
                        if (f < s1) {
                            int r = 0;
#pragma unroll
                            for (int st = 8; st > 0; st >>= 1)
                                if (tpre[r + st] <= f) r += st;
                            const int p = tstart[r] + (f - tpre[r]);
                            rr[u] = r;
                            term[u] = indices[p];
                            val[u] = values[p];
                        }
                    }
                    int2 col[kA];
#pragma unroll
                    for (int u = 0; u < kA; ++u) col[u] = rr[u] >= 0 ? tmap[term[u]] : make_int2(-1, -1);
#pragma unroll
                    for (int u = 0; u < kA; ++u) {
                        const int r = rr[u];
                        const int2 cl = col[u];
                        if (r >= 0 && cl.x <= -2) {
                            atile[r * kHot + (-2 - cl.x)] = static_cast<__bf16>(val[u]);
                            n_pairs += static_cast<uint64_t>(cl.y);
                        }
                        const bool cold = r >= 0 && cl.x >= 0 && cl.y > cl.x;
                        const uint64_t m = __ballot(cold);
                        int base = 0;
                        if (lane == 0 && m != 0) base = atomicAdd(&misc[2], __popcll(m));
                        base = __shfl(base, 0, 64);
                        if (cold) {
                            const int at = base + __popcll(m & ((1ULL << lane) - 1));
                            ccs[at] = cl.x;
                            cpre[at] = cl.y - cl.x;
                            crd[at] = r | (val[u] << 8);
                        }
                    }
                }
                __syncthreads();
                // (2) the cold postings: exclusive prefix of the counts (four
                // entries per lane, then across lanes and waves), then pairs
                const int M = misc[2];
                int loc[4], lsum = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = tid * 4 + u;
                    loc[u] = i < M ? cpre[i] : 0;
                    lsum += loc[u];
                }
                const int wincl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(lsum)));
                int* wtot = misc + 4;  // [kWaves]
                if (lane == 63) wtot[wave] = wincl;
                __syncthreads();
                int wbase = 0;
                for (int w = 0; w < wave; ++w) wbase += wtot[w];
                int run = wbase + wincl - lsum;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = tid * 4 + u;
                    if (i < M) cpre[i] = run;
                    run += loc[u];
                }
                if (tid == kHB - 1) cpre[M] = run;  // every count is in by the last lane
                __syncthreads();
                const int P = cpre[M];
                constexpr int kU = 8;
                for (int i0 = 0; i0 < P; i0 += kHB * kU) {
                    int ent[kU], off[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int i = i0 + u * kHB + tid;
                        int lo = 0, hi = M;  // the last entry whose prefix is <= i
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (cpre[mid] <= i)
                                lo = mid;
                            else
                                hi = mid;
                        }
                        ent[u] = i < P ? lo : -1;
                        off[u] = i - cpre[lo];
                    }
                    uint32_t cl[kU], cwv[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int64_t e = ent[u] >= 0 ? static_cast<int64_t>(ccs[ent[u]]) + off[u] : 0;
                        cl[u] = ent[u] >= 0 ? cc[e] : 0u;
                        cwv[u] = ent[u] >= 0 ? static_cast<uint32_t>(cw[e]) : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if (ent[u] < 0) continue;
                        const int rd = crd[ent[u]];
                        atomicAdd(&sims[(rd & 0xFF) * nc + static_cast<int>(cl[u])],
                                  (rd >> 8) * static_cast<int32_t>(cwv[u]));
                    }
                }
                n_pairs += static_cast<uint64_t>(tid == 0 ? P : 0);
                __syncthreads();
            }
            // (3) dense part: 16 docs x 16 centroids per MFMA tile, K = 32 hot
            // slots per step; lane l holds A[l % 16][8 (l / 16) ..],
            // B[8 (l / 16) ..][l % 16] and D[4 (l / 16) + v][l % 16]. All K
            // steps' B fragments are loaded before the first MFMA (slots past
            // the hot ones are zero in both operands).
            if (nhot > 0) {
                const int arow = lane & 15, kq = 8 * (lane >> 4);
                for (int ct = wave; ct * 16 < nc; ct += kWaves) {
                    const int c = ct * 16 + (lane & 15);
                    const __bf16* brow = bmat + static_cast<size_t>(c) * kHot + kq;
                    bf16x8_t bv[kHot / 32];
#pragma unroll
                    for (int ks = 0; ks < kHot / 32; ++ks) bv[ks] = *reinterpret_cast<const bf16x8_t*>(brow + ks * 32);
                    f32x4_t acc = {0.0F, 0.0F, 0.0F, 0.0F};
#pragma unroll
                    for (int ks = 0; ks < kHot / 32; ++ks) {
                        const bf16x8_t av = *reinterpret_cast<const bf16x8_t*>(atile + arow * kHot + ks * 32 + kq);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[ks], acc, 0, 0, 0);
                    }
                    if (c < nc) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int r = 4 * (lane >> 4) + v;
                            if (r < nt) sims[r * nc + c] += static_cast<int32_t>(acc[v]);
                        }
                    }
                }
            }
            __syncthreads();
            // (4) argmax per doc (strict >: the lowest cluster wins ties)
            for (int r = wave; r < nt; r += kWaves) {
                if (tdoc[r] < 0) {
                    if (lane == 0) assign[kb + j0 + r] = -1;  // a centroid
                    continue;
                }
                const int32_t* srow = sims + r * nc;
                int32_t best = srow[0];
                int bi = 0;
                for (int c = lane; c < nc; c += 64) argmax_merge(best, bi, srow[c], c);
                for (int off = 32; off > 0; off >>= 1) {
                    const int32_t ov = __shfl_xor(best, off, 64);
                    const int oi = __shfl_xor(bi, off, 64);
                    argmax_merge(best, bi, ov, oi);
                }
                if (lane == 0) assign[kb + j0 + r] = bi;
            }
            __syncthreads();
        }
        for (int64_t e = e0 + tid; e < e1; e += kHB) tmap[ckey[e] & 0xFFFFu] = make_int2(-1, -1);
        __syncthreads();
    }
    // the lanes' pair counts, summed per wave
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor(static_cast<uint32_t>(n_pairs), off, 64);
        const uint32_t hi = __shfl_xor(static_cast<uint32_t>(n_pairs >> 32), off, 64);
        n_pairs += (static_cast<uint64_t>(hi) << 32) | lo;
    }
    if (lane == 0 && n_pairs != 0) atomicAdd(pairs, static_cast<unsigned long long>(n_pairs));
}

// Non-zeros of the kept postings' rows (for the build's byte count).
__global__ void kept_row_nnz_kernel(const int32_t* __restrict__ indptr, const int32_t* __restrict__ kept, size_t n,
                                    unsigned long long* __restrict__ out) {
    uint64_t local = 0;
    GRID_STRIDE(i, n) local += static_cast<uint64_t>(indptr[kept[i] + 1] - indptr[kept[i]]);
    for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
    if ((threadIdx.x & 63) == 0 && local != 0) atomicAdd(out, static_cast<unsigned long long>(local));
}

// Cluster sizes: centroid + assigned docs (per global cluster).
__global__ void cluster_count_kernel(const int64_t* __restrict__ kept_base, const int32_t* __restrict__ assign,
                                     const int32_t* __restrict__ lcb, int64_t* __restrict__ count, int n_lists) {
    for (int l = blockIdx.x; l < n_lists; l += gridDim.x) {
        const int c0 = lcb[l], nc = lcb[l + 1] - c0;
        for (int c = threadIdx.x; c < nc; c += blockDim.x) count[c0 + c] = 1;
        __syncthreads();
        const int64_t kb = kept_base[l], n = kept_base[l + 1] - kb;
        for (int64_t j = threadIdx.x; j < n; j += blockDim.x) {
            const int a = assign[kb + j];
            if (a >= 0) atomicAdd(reinterpret_cast<unsigned long long*>(&count[c0 + a]), 1ULL);
        }
        __syncthreads();
    }
}

// Stable placement: cluster c = [centroid_c, its docs in pruned order].
__global__ __launch_bounds__(64) void cluster_place_kernel(const int64_t* __restrict__ kept_base,
                                                           const int32_t* __restrict__ kept,
                                                           const int32_t* __restrict__ assign,
                                                           const int32_t* __restrict__ lcb,
                                                           const int32_t* __restrict__ centroid_doc,
                                                           const int64_t* __restrict__ cdp, int32_t* __restrict__ docs,
                                                           int n_lists) {
    __shared__ int32_t fill[4096];
    const int lane = threadIdx.x;
    for (int l = blockIdx.x; l < n_lists; l += gridDim.x) {
        const int c0 = lcb[l], nc = lcb[l + 1] - c0;
        for (int c = lane; c < nc; c += 64) {
            docs[cdp[c0 + c]] = centroid_doc[c0 + c];
            fill[c] = 1;
        }
        __builtin_amdgcn_wave_barrier();
        const int64_t kb = kept_base[l], n = kept_base[l + 1] - kb;
        for (int64_t b = 0; b < n; b += 64) {
            const int64_t j = b + lane;
            const int a = j < n ? assign[kb + j] : -1;
            int before = 0, after = 0;
            for (int o = 0; o < 64; ++o) {
                const int oa = __shfl(a, o, 64);
                if (a >= 0 && oa == a) {
                    if (o < lane) ++before;
                    if (o > lane) ++after;
                }
            }
            int64_t base = 0;
            if (a >= 0) base = fill[a];
            __builtin_amdgcn_wave_barrier();
            if (a >= 0) {
                docs[cdp[c0 + a] + base + before] = kept[kb + j];
                if (after == 0) fill[a] = static_cast<int32_t>(base + before + 1);
            }
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- uint8 summaries: summarize_with_cpu_<uint8_t> + summarize_emit_cluster_
// (cluster/inverted_list_clusters.cpp:37-124). With integer values the running
// float sum is exact, so the sum is the sum of per-term maxima; the alpha cut
// is found from a 256-bin histogram of the maxima, and ties at the cut are
// kept in term order.
struct SumLds8 {
    uint32_t* maxw;   // [nw*8] four uint8 maxima per word
    uint32_t* touch;  // [nw]
    int32_t* hist;    // [256]
    int32_t* wscan;   // [nw] per-word kept counts, scanned
    int32_t* wtie;    // [nw] cut-valued terms before each word
    int32_t* misc;    // [8]
};

__device__ __forceinline__ void byte_max(uint32_t* w, int t, uint32_t v) {
    uint32_t* word = w + (t >> 2);
    const int sh = (t & 3) * 8;
    uint32_t old = *word;
    while (((old >> sh) & 0xFFu) < v) {
        const uint32_t nw = (old & ~(0xFFu << sh)) | (v << sh);
        const uint32_t prev = atomicCAS(word, old, nw);
        if (prev == old) break;
        old = prev;
    }
}

__device__ __forceinline__ uint32_t max_of(const uint32_t* w, int t) { return (w[t >> 2] >> ((t & 3) * 8)) & 0xFFu; }

// The alpha cut of one cluster by one wavefront, from the 256-bin histogram of
// its per-term maxima and their exact integer sum. The reference sorts the
// (term, max) pairs by value desc, adds them up in that order and stops at the
// first addup / sum >= alpha (inverted_list_clusters.cpp:43-50). Lane l holds
// values 255 - 4l - u (u < 4); a DPP prefix gives each value the exact sum of
// all higher ones (a), and every non-empty value runs the sequential loop's
// test float(a + v * m) / sum >= alpha; the first value that passes (in
// descending order) is the cut, and its lane finds how many of its equal terms
// are kept. Returns (cut value, terms kept at it), the same in every lane;
// (-1, 0) when no value passes: every touched term is kept.
__device__ int2 alpha_cut(const int32_t* hist, int sum_i, float alpha, int lane) {
    const float sum = static_cast<float>(sum_i);
    int m4[4];
    uint32_t part = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        m4[u] = hist[255 - 4 * lane - u];
        part += static_cast<uint32_t>((255 - 4 * lane - u) * m4[u]);
    }
    int64_t a = static_cast<int64_t>(wave_incl_scan_dpp(part) - part);
    int hit = -1;
    int64_t a_hit = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int v = 255 - 4 * lane - u, m = m4[u];
        if (hit < 0 && m != 0 && static_cast<float>(a + static_cast<int64_t>(v) * m) / sum >= alpha) {
            hit = u;
            a_hit = a;
        }
        a += static_cast<int64_t>(v) * m;
    }
    const uint64_t hm = __ballot(hit >= 0);
    if (hm == 0) return make_int2(-1, 0);
    const int first = __ffsll(static_cast<unsigned long long>(hm)) - 1;
    int cv = 0, lo = 1;
    if (lane == first) {
        const int v = 255 - 4 * lane - hit, m = m4[hit];
        int hi = m;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (static_cast<float>(a_hit + static_cast<int64_t>(v) * mid) / sum >= alpha)
                hi = mid;
            else
                lo = mid + 1;
        }
        cv = v;
    }
    return make_int2(__shfl(cv, first, 64), __shfl(lo, first, 64));
}

// Clusters whose rows hold at most kSumSmallCap entries are summarized by one
// wavefront each: the (term << 8 | code) entries of the cluster's rows are
// bitonic-sorted in LDS, so each term's last entry carries its maximum and the
// terms come in ascending order. Then come the histogram of the maxima, the
// alpha cut, and the kept terms -- those above the cut value, and the first
// cut_x at it in term order -- counted (count pass) or written at out_off[g]
// (write pass), exactly as summarize_u8_kernel does. Larger clusters are left
// to it: the count pass appends them to `big`.
constexpr int kSumSmallCap = 1024;
constexpr int kSumWaves = kB / 64;

template <bool WRITE>
__global__ __launch_bounds__(kB) void summarize_small_u8_kernel(
    const int32_t* __restrict__ indptr, const uint16_t* __restrict__ indices, const uint8_t* __restrict__ values,
    const int64_t* __restrict__ cdp, const int32_t* __restrict__ docs, const int32_t* __restrict__ cluster_list,
    const int32_t* __restrict__ lcb, int64_t n_clusters, float alpha, int64_t* __restrict__ count,
    const int64_t* __restrict__ out_off, uint16_t* __restrict__ out_term, uint8_t* __restrict__ out_val,
    uint16_t* __restrict__ out_cl, int32_t* __restrict__ big, int32_t* __restrict__ n_big) {
    __shared__ uint32_t ent_all[kSumWaves][kSumSmallCap];
    __shared__ int32_t hist_all[kSumWaves][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t* ent = ent_all[wave];
    int32_t* hist = hist_all[wave];
    const uint64_t below = (uint64_t{1} << lane) - 1;
    for (int64_t g = static_cast<int64_t>(blockIdx.x) * kSumWaves + wave; g < n_clusters;
         g += static_cast<int64_t>(gridDim.x) * kSumWaves) {
        const int64_t q0 = cdp[g], q1 = cdp[g + 1];
        int n_e = 0;
        for (int64_t q = q0 + lane; q < q1; q += 64) {
            const int32_t d = docs[q];
            n_e += indptr[d + 1] - indptr[d];
        }
        for (int off = 32; off > 0; off >>= 1) n_e += __shfl_xor(n_e, off, 64);
        if (n_e > kSumSmallCap) {
            if (!WRITE && lane == 0) big[atomicAdd(n_big, 1)] = static_cast<int32_t>(g);
            continue;
        }
        int pos = 0;
        for (int64_t q = q0; q < q1; ++q) {
            const int32_t d = docs[q], r0 = indptr[d], r1 = indptr[d + 1];
            for (int32_t p = r0 + lane; p < r1; p += 64)
                ent[pos + (p - r0)] = (static_cast<uint32_t>(indices[p]) << 8) | values[p];
            pos += r1 - r0;
        }
        int n2 = 64;
        while (n2 < n_e) n2 <<= 1;
        for (int i = n_e + lane; i < n2; i += 64) ent[i] = 0xFFFFFFFFu;
        for (int i = lane; i < 256; i += 64) hist[i] = 0;
        __builtin_amdgcn_wave_barrier();
        for (int size = 2; size <= n2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = lane; t < n2 / 2; t += 64) {
                    const int i = 2 * t - (t & (stride - 1)), j = i + stride;
                    const uint32_t a = ent[i], b = ent[j];
                    if ((a > b) == ((i & size) == 0)) {
                        ent[i] = b;
                        ent[j] = a;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        int local_sum = 0;
        for (int i = lane; i < n_e; i += 64) {
            const uint32_t k = ent[i];
            if (i + 1 == n_e || (ent[i + 1] >> 8) != (k >> 8)) {
                atomicAdd(&hist[k & 0xFFu], 1);
                local_sum += static_cast<int>(k & 0xFFu);
            }
        }
        for (int off = 32; off > 0; off >>= 1) local_sum += __shfl_xor(local_sum, off, 64);
        __builtin_amdgcn_wave_barrier();
        const int2 cut = alpha_cut(hist, local_sum, alpha, lane);
        int run_tie = 0, run_keep = 0;
        const int64_t base = WRITE ? out_off[g] : 0;
        const uint16_t local_c = WRITE ? static_cast<uint16_t>(g - lcb[cluster_list[g]]) : uint16_t{0};
        for (int i0 = 0; i0 < n_e; i0 += 64) {
            const int i = i0 + lane;
            uint32_t k = 0;
            bool last = false;
            if (i < n_e) {
                k = ent[i];
                last = i + 1 == n_e || (ent[i + 1] >> 8) != (k >> 8);
            }
            const int v = static_cast<int>(k & 0xFFu);
            const bool tie = last && v == cut.x;
            const uint64_t tb = __ballot(tie);
            const int tie_rank = run_tie + __popcll(tb & below);
            const bool keep = last && (v > cut.x || (tie && tie_rank < cut.y));
            const uint64_t kb = __ballot(keep);
            if (WRITE && keep) {
                const int64_t o = base + run_keep + __popcll(kb & below);
                out_term[o] = static_cast<uint16_t>(k >> 8);
                out_val[o] = static_cast<uint8_t>(v);
                out_cl[o] = local_c;
            }
            run_tie += __popcll(tb);
            run_keep += __popcll(kb);
        }
        if (!WRITE && lane == 0) count[g] = run_keep;
        __builtin_amdgcn_wave_barrier();
    }
}

template <bool WRITE>
__global__ __launch_bounds__(kB) void summarize_u8_kernel(const int32_t* __restrict__ indptr,
                                                          const uint16_t* __restrict__ indices,
                                                          const uint8_t* __restrict__ values,
                                                          const int64_t* __restrict__ cdp,
                                                          const int32_t* __restrict__ docs, const int32_t* __restrict__ cluster_list,
                                                          const int32_t* __restrict__ lcb, int64_t n_clusters, int nw,
                                                          float alpha, int64_t* __restrict__ count,
                                                          const int64_t* __restrict__ out_off,
                                                          uint16_t* __restrict__ out_term, uint8_t* __restrict__ out_val,
                                                          uint16_t* __restrict__ out_cl,
                                                          const int32_t* __restrict__ list,
                                                          const int32_t* __restrict__ n_list) {
    extern __shared__ __align__(16) unsigned char smem[];
    SumLds8 S;
    S.maxw = reinterpret_cast<uint32_t*>(smem);
    S.touch = S.maxw + nw * 8;
    S.hist = reinterpret_cast<int32_t*>(S.touch + nw);
    S.wscan = S.hist + 256;
    S.wtie = S.wscan + nw;
    S.misc = S.wtie + nw;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int w = tid; w < nw * 8; w += kB) S.maxw[w] = 0;
    for (int w = tid; w < nw; w += kB) S.touch[w] = 0;
    __syncthreads();
    // every cluster, or those of `list` (n_list[0] of them)
    const int64_t n_items = list != nullptr ? *n_list : n_clusters;
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int64_t g = list != nullptr ? list[it] : it;
        for (int i = tid; i < 256; i += kB) S.hist[i] = 0;
        // max-pool over the cluster's docs
        for (int64_t q = cdp[g] + wave; q < cdp[g + 1]; q += kB / 64) {
            const int32_t d = docs[q];
            for (int32_t p = indptr[d] + lane; p < indptr[d + 1]; p += 64) {
                const int t = indices[p];
                atomicOr(&S.touch[t >> 5], 1u << (t & 31));
                byte_max(S.maxw, t, values[p]);
            }
        }
        __syncthreads();
        // histogram of maxima, exact integer sum
        int local_sum = 0;
        for (int w = tid; w < nw; w += kB) {
            uint32_t bits = S.touch[w];
            while (bits) {
                const int b = __ffs(bits) - 1;
                bits &= bits - 1;
                const uint32_t v = max_of(S.maxw, w * 32 + b);
                atomicAdd(&S.hist[v], 1);
                local_sum += static_cast<int>(v);
            }
        }
        if (tid == 0) S.misc[0] = 0;
        __syncthreads();
        atomicAdd(&S.misc[0], local_sum);
        __syncthreads();
        if (wave == 0) {
            const int2 cut = alpha_cut(S.hist, S.misc[0], alpha, lane);
            if (lane == 0) {
                S.misc[1] = cut.x;
                S.misc[2] = cut.y;
            }
        }
        __syncthreads();
        const int cut_v = S.misc[1], cut_x = S.misc[2];
        // rank of each cut-valued term in term order -> per-word counts, scan
        // (single wave, word order), then per-word kept counts and output.
        if (wave == 0) {
            int run_tie = 0, run_keep = 0;
            for (int b = 0; b < nw; b += 64) {
                const int w = b + lane;
                int ties = 0;
                uint32_t bits = w < nw ? S.touch[w] : 0u;
                uint32_t tb = bits;
                while (tb) {
                    const int bb = __ffs(tb) - 1;
                    tb &= tb - 1;
                    if (static_cast<int>(max_of(S.maxw, w * 32 + bb)) == cut_v) ++ties;
                }
                const int incl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(ties)));
                int tie_rank = run_tie + incl - ties;
                int keep = 0;
                tb = bits;
                while (tb) {
                    const int bb = __ffs(tb) - 1;
                    tb &= tb - 1;
                    const int v = static_cast<int>(max_of(S.maxw, w * 32 + bb));
                    if (v > cut_v)
                        ++keep;
                    else if (v == cut_v) {
                        if (tie_rank < cut_x || cut_v < 0) ++keep;
                        ++tie_rank;
                    }
                }
                const int kincl = static_cast<int>(wave_incl_scan_dpp(static_cast<uint32_t>(keep)));
                if (w < nw) {
                    S.wscan[w] = run_keep + kincl - keep;
                    S.wtie[w] = run_tie + incl - ties;
                }
                run_tie += __builtin_amdgcn_readlane(incl, 63);
                run_keep += __builtin_amdgcn_readlane(kincl, 63);
            }
            if (lane == 0) S.misc[3] = run_keep;
        }
        __syncthreads();
        if (!WRITE) {
            if (tid == 0) count[g] = S.misc[3];
        } else {
            const int64_t base = out_off[g];
            const uint16_t local_c = static_cast<uint16_t>(g - lcb[cluster_list[g]]);
            // recompute tie ranks per word from the scanned tie counts
            for (int w = tid; w < nw; w += kB) {
                uint32_t bits = S.touch[w];
                if (!bits) continue;
                int pos = S.wscan[w];
                int tie_before = S.wtie[w];
                while (bits) {
                    const int bb = __ffs(bits) - 1;
                    bits &= bits - 1;
                    const int t = w * 32 + bb;
                    const int v = static_cast<int>(max_of(S.maxw, t));
                    bool keep = v > cut_v;
                    if (v == cut_v) {
                        keep = cut_v < 0 || tie_before < cut_x;
                        ++tie_before;
                    }
                    if (keep) {
                        out_term[base + pos] = static_cast<uint16_t>(t);
                        out_val[base + pos] = static_cast<uint8_t>(v);
                        out_cl[base + pos] = local_c;
                        ++pos;
                    }
                }
            }
        }
        __syncthreads();
        // clear the touched words for the next cluster
        for (int w = tid; w < nw; w += kB) {
            if (S.touch[w]) {
                S.touch[w] = 0;
                for (int j = 0; j < 8; ++j) S.maxw[w * 8 + j] = 0;
            }
        }
        __syncthreads();
    }
}

// ---- FP32 and 16-bit summaries: the reference's sequential algorithm (one
// lane walks the cluster's events in order so the running float sum rounds
// exactly as the CPU's does), an LDS dense accumulator, and a workgroup bitonic
// sort of the (max desc, term asc) pairs -- in LDS when they fit, else in this
// workgroup's global scratch (`spill`, next_pow2(dim) entries). 16-bit codes
// sit in the float accumulator exactly, and the sum adds the same floats as the
// reference's (float += code, float += integer increase).
template <class T>
__global__ __launch_bounds__(kB) void summarize_seq_kernel(const int32_t* __restrict__ indptr,
                                                           const uint16_t* __restrict__ indices,
                                                           const T* __restrict__ values,
                                                           const int64_t* __restrict__ cdp,
                                                           const int32_t* __restrict__ docs,
                                                           const int32_t* __restrict__ cluster_list,
                                                           const int32_t* __restrict__ lcb, int64_t n_clusters, int dim,
                                                           int cap, float alpha, int64_t* __restrict__ count,
                                                           const int64_t* __restrict__ out_off, bool write,
                                                           uint16_t* __restrict__ out_term, T* __restrict__ out_val,
                                                           uint16_t* __restrict__ out_cl, KeyItem* __restrict__ spill,
                                                           int spill_stride) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nw = (dim + 31) / 32;
    float* acc = reinterpret_cast<float*>(smem);
    uint32_t* touch = reinterpret_cast<uint32_t*>(acc + dim);
    uint32_t* keepb = touch + nw;
    const size_t pairs_at = (static_cast<size_t>(dim) * 4 + static_cast<size_t>(nw) * 8 + 7) & ~size_t{7};
    KeyItem* lds_pairs = reinterpret_cast<KeyItem*>(smem + pairs_at);
    int32_t* misc = reinterpret_cast<int32_t*>(lds_pairs + cap);
    int32_t* wsum = misc + 4;
    KeyItem* my_spill = spill + static_cast<size_t>(blockIdx.x) * spill_stride;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int w = tid; w < nw; w += kB) {
        touch[w] = 0;
        keepb[w] = 0;
    }
    __syncthreads();
    for (int64_t g = blockIdx.x; g < n_clusters; g += gridDim.x) {
        if (tid == 0) {
            float sum = 0.0F;
            int m = 0;
            for (int64_t q = cdp[g]; q < cdp[g + 1]; ++q) {
                const int32_t d = docs[q];
                for (int32_t p = indptr[d]; p < indptr[d + 1]; ++p) {
                    const int t = indices[p];
                    const float v = static_cast<float>(values[p]);
                    if (!((touch[t >> 5] >> (t & 31)) & 1u)) {
                        touch[t >> 5] |= 1u << (t & 31);
                        const float value = (0.0F < v) ? v : 0.0F;  // std::max(0.0f, v)
                        acc[t] = value;
                        sum += value;
                        ++m;
                    } else if (v > acc[t]) {
                        sum += v - acc[t];
                        acc[t] = v;
                    }
                }
            }
            misc[0] = m;
            reinterpret_cast<float*>(misc)[1] = sum;
        }
        __syncthreads();
        const int m = misc[0];
        const float sum = reinterpret_cast<float*>(misc)[1];
        int np = 1;
        while (np < m) np <<= 1;
        KeyItem* pairs = np <= cap ? lds_pairs : my_spill;
        // touched terms -> pairs (positions from a popcount scan over the words)
        {
            const int per = (nw + kB - 1) / kB;
            const int w0 = tid * per;
            int local = 0;
            for (int w = w0; w < w0 + per && w < nw; ++w) local += __popc(touch[w]);
            int incl = local;
            for (int o = 1; o < 64; o <<= 1) {
                const int x = __shfl_up(incl, o, 64);
                if (lane >= o) incl += x;
            }
            if (lane == 63) wsum[wave] = incl;
            __syncthreads();
            int run = incl - local;
            for (int w = 0; w < wave; ++w) run += wsum[w];
            for (int w = w0; w < w0 + per && w < nw; ++w) {
                uint32_t bits = touch[w];
                while (bits) {
                    const int b = __ffs(bits) - 1;
                    bits &= bits - 1;
                    const int t = w * 32 + b;
                    pairs[run++] = KeyItem{acc[t], static_cast<uint32_t>(t)};
                }
            }
        }
        for (int i = m + tid; i < np; i += kB) pairs[i] = KeyItem{-__builtin_inff(), 0xFFFFFFFFu};
        __syncthreads();
        for (int size = 2; size <= np; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < np; i += kB) {
                    const int j = i ^ stride;
                    if (j > i) {
                        const KeyItem a = pairs[i], b = pairs[j];
                        const bool a_first = a.key > b.key || (a.key == b.key && a.id < b.id);
                        const bool up = (i & size) == 0;
                        if (up ? !a_first : a_first) {
                            pairs[i] = b;
                            pairs[j] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (tid == 0) {
            // summarize_emit_cluster_ (inverted_list_clusters.cpp:43-50)
            float addup = 0.0F;
            int keep_n = m;
            for (int j = 0; j < m; ++j) {
                addup += pairs[j].key;
                if (addup / sum >= alpha) {
                    keep_n = j + 1;
                    break;
                }
            }
            misc[2] = keep_n;
        }
        __syncthreads();
        const int keep_n = misc[2];
        for (int i = tid; i < keep_n; i += kB) atomicOr(&keepb[pairs[i].id >> 5], 1u << (pairs[i].id & 31));
        __syncthreads();
        if (!write) {
            if (tid == 0) count[g] = keep_n;
        } else if (tid == 0) {
            int64_t pos = out_off[g];
            const uint16_t local_c = static_cast<uint16_t>(g - lcb[cluster_list[g]]);
            for (int w = 0; w < nw; ++w) {
                uint32_t bits = keepb[w];
                while (bits) {
                    const int b = __ffs(bits) - 1;
                    bits &= bits - 1;
                    const int t = w * 32 + b;
                    out_term[pos] = static_cast<uint16_t>(t);
                    out_val[pos] = static_cast<T>(acc[t]);
                    out_cl[pos] = local_c;
                    ++pos;
                }
            }
        }
        __syncthreads();
        for (int w = tid; w < nw; w += kB) {
            touch[w] = 0;
            keepb[w] = 0;
        }
        __syncthreads();
    }
}

// Transpose keys (list, term) for clusters [g0, g0 + n_clusters); entry ids
// relative to e_base, the first summary entry of cluster g0.
__global__ void summary_key_kernel(const uint16_t* __restrict__ term, const int64_t* __restrict__ out_off,
                                   const int32_t* __restrict__ cluster_list, size_t g0, size_t n_clusters,
                                   int64_t e_base, uint32_t* __restrict__ key, int32_t* __restrict__ val) {
    GRID_STRIDE(i, n_clusters) {
        const size_t g = g0 + i;
        const uint32_t l = static_cast<uint32_t>(cluster_list[g]);
        for (int64_t e = out_off[g]; e < out_off[g + 1]; ++e) {
            key[e - e_base] = (l << 16) | term[e];
            val[e - e_base] = static_cast<int32_t>(e - e_base);
        }
    }
}

template <class T>
__global__ void gather_summary_kernel(const int32_t* __restrict__ perm, const uint16_t* __restrict__ cl,
                                      const T* __restrict__ v, size_t n, uint16_t* __restrict__ cl_out,
                                      T* __restrict__ v_out) {
    GRID_STRIDE(i, n) {
        cl_out[i] = cl[perm[i]];
        v_out[i] = v[perm[i]];
    }
}

__global__ void head_flag_kernel(const uint32_t* __restrict__ key, size_t n, int64_t* __restrict__ flag) {
    GRID_STRIDE(i, n) flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// One batch of the transpose: batch entry i is global entry e_base + i; its
// head ordinal within the batch is hscan[i].
__global__ void head_write_kernel(const uint32_t* __restrict__ key, const int64_t* __restrict__ hscan, size_t n,
                                  int64_t e_base, uint16_t* __restrict__ term_ids, int64_t* __restrict__ term_ptr) {
    GRID_STRIDE(i, n) {
        if (i == 0 || key[i] != key[i - 1]) {
            const int64_t h = hscan[i];
            term_ids[h] = static_cast<uint16_t>(key[i] & 0xFFFFu);
            term_ptr[h] = e_base + static_cast<int64_t>(i);
        }
    }
}

// Per-list entry and term bases for lists [l0, l1), whose summary entries are
// the batch [e_base, e_base + n) with batch head scan `hscan` (n + 1 values);
// h_base is the number of term heads in earlier batches.
__global__ void list_bases_kernel(const int32_t* __restrict__ lcb, const int64_t* __restrict__ out_off,
                                  const int64_t* __restrict__ hscan, int64_t e_base, int64_t n, int64_t h_base,
                                  int l0, int l1, int64_t* __restrict__ leb, int32_t* __restrict__ ltb) {
    GRID_STRIDE(i, static_cast<size_t>(l1 - l0)) {
        const int l = l0 + static_cast<int>(i);
        const int64_t e = out_off[lcb[l]];
        leb[l] = e;
        ltb[l] = static_cast<int32_t>(h_base + hscan[min(e - e_base, n)]);
    }
}

// ---------------------------------------------------------------- block rows

// Encodes one uint8 row into block words (the layout in device_index.h) at
// `out` and returns their count; with out == nullptr it only counts. One
// thread per row: the low terms first, 8 slots to a word, then the other
// terms in row order under word bases.
__device__ int encode_block_row(const uint16_t* __restrict__ t, const uint8_t* __restrict__ c, int len,
                                uint4* __restrict__ out) {
    int nw = 0, ns = 0;
    uint64_t lo = 0, hi = 0;
    auto put = [&](uint32_t slot) {
        if (ns < 4)
            lo |= static_cast<uint64_t>(slot) << (16 * ns);
        else
            hi |= static_cast<uint64_t>(slot) << (16 * (ns - 4));
        ++ns;
    };
    auto flush = [&]() {
        if (out != nullptr)
            out[nw] = make_uint4(static_cast<uint32_t>(lo), static_cast<uint32_t>(lo >> 32),
                                 static_cast<uint32_t>(hi), static_cast<uint32_t>(hi >> 32));
        ++nw;
        ns = 0;
        lo = hi = 0;
    };
    for (int i = 0; i < len; ++i) {
        const uint32_t term = t[i];
        uint32_t code = c[i];
        if (term >= static_cast<uint32_t>(kBlkTermSpan)) continue;
        while (code > 0) {
            const uint32_t part = min(code, 15U);
            put((term << 4) | part);
            code -= part;
            if (ns == 8) flush();
        }
    }
    if (ns > 0) flush();
    uint32_t base = 0;
    for (int i = 0; i < len; ++i) {
        const uint32_t term = t[i];
        uint32_t code = c[i];
        if (term < static_cast<uint32_t>(kBlkTermSpan)) continue;
        while (code > 0) {
            if (ns > 0 && (term < base || term - base >= static_cast<uint32_t>(kBlkTermSpan))) flush();
            if (ns == 0) {
                base = term & ~15U;
                put(base);  // slot 0: the base, code nibble 0
            }
            const uint32_t part = min(code, 15U);
            put(((term - base) << 4) | part);
            code -= part;
            if (ns == 8) flush();
        }
    }
    if (ns > 0) flush();
    return nw;
}

// Block words of each doc's encoded row, rounded up to whole pairs (the
// search reads a row as 32-byte pairs of words); flags rows whose nnz or word
// count exceed the row_ref fields.
__global__ void doc_block_words_kernel(const int32_t* __restrict__ indptr, const uint16_t* __restrict__ indices,
                                       const uint8_t* __restrict__ values, size_t n, int64_t* __restrict__ words,
                                       int32_t* __restrict__ too_long) {
    GRID_STRIDE(d, n) {
        const int32_t r0 = indptr[d], len = indptr[d + 1] - r0;
        const int nw = (encode_block_row(indices + r0, values + r0, len, nullptr) + 1) & ~1;
        words[d] = nw;
        if (len > kBlkMaxField || nw > kBlkMaxField) atomicOr(too_long, 1);
    }
}

// Each doc's encoded row at its offset dws[d] of `rows`, a zero word
// completing an odd count.
__global__ void doc_block_encode_kernel(const int32_t* __restrict__ indptr, const uint16_t* __restrict__ indices,
                                        const uint8_t* __restrict__ values, size_t n, const int64_t* __restrict__ dws,
                                        uint4* __restrict__ rows) {
    GRID_STRIDE(d, n) {
        const int32_t r0 = indptr[d], len = indptr[d + 1] - r0;
        const int nw = encode_block_row(indices + r0, values + r0, len, rows + dws[d]);
        if (nw & 1) rows[dws[d] + nw] = make_uint4(0, 0, 0, 0);
    }
}

// Word count of each posting's row copy.
__global__ void block_words_kernel(const int32_t* __restrict__ docs, const int64_t* __restrict__ dws, size_t n,
                                   int64_t* __restrict__ words) {
    GRID_STRIDE(p, n) {
        const int32_t d = docs[p];
        words[p] = dws[d + 1] - dws[d];
    }
}

// Largest word span of one posting list's row copies.
__global__ void list_word_span_kernel(const int32_t* __restrict__ lcb, const int64_t* __restrict__ cdp,
                                      const int64_t* __restrict__ ws, int n_lists,
                                      unsigned long long* __restrict__ max_span) {
    GRID_STRIDE(l, static_cast<size_t>(n_lists)) {
        const int64_t a = ws[cdp[lcb[l]]], b = ws[cdp[lcb[l + 1]]];
        atomicMax(max_span, static_cast<unsigned long long>(b - a));
    }
}

// One 16-lane group per posting copies its doc's encoded words (16 B each)
// and writes row_ref.
__global__ __launch_bounds__(kB) void block_copy_kernel(const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ docs,
                                                        const int64_t* __restrict__ dws,
                                                        const uint4* __restrict__ doc_rows,
                                                        const int64_t* __restrict__ ws, size_t n,
                                                        uint4* __restrict__ blk_rows,
                                                        int64_t* __restrict__ row_ref) {
    const int g = threadIdx.x >> 4, gl = threadIdx.x & 15;
    for (size_t p = static_cast<size_t>(blockIdx.x) * (kB / 16) + g; p < n;
         p += static_cast<size_t>(gridDim.x) * (kB / 16)) {
        const int32_t d = docs[p];
        const int64_t s0 = dws[d];
        const int nw = static_cast<int>(dws[d + 1] - s0);
        const int64_t w0 = ws[p];
        if (gl == 0) row_ref[p] = blk_ref(w0, nw, indptr[d + 1] - indptr[d]);
        for (int w = gl; w < nw; w += 16) blk_rows[w0 + w] = doc_rows[s0 + w];
    }
}

// One workgroup per posting list: its summary terms as a bitmap over the
// vocabulary (in LDS, then written out) and the exclusive prefix of the
// words' set bits.
__global__ __launch_bounds__(kB) void summary_bitmap_kernel(const int32_t* __restrict__ ltb,
                                                            const uint16_t* __restrict__ tids, int nw,
                                                            uint32_t* __restrict__ bits, uint16_t* __restrict__ rank) {
    extern __shared__ uint32_t sbits[];  // nw words, then kB / 64 wave sums
    int32_t* wsum = reinterpret_cast<int32_t*>(sbits + nw);
    const int l = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int w = tid; w < nw; w += kB) sbits[w] = 0;
    __syncthreads();
    for (int32_t i = ltb[l] + tid; i < ltb[l + 1]; i += kB) atomicOr(&sbits[tids[i] >> 5], 1u << (tids[i] & 31));
    __syncthreads();
    const int per = (nw + kB - 1) / kB, w0 = tid * per;
    int local = 0;
    for (int w = w0; w < w0 + per && w < nw; ++w) local += __popc(sbits[w]);
    int incl = local;
    for (int o = 1; o < 64; o <<= 1) {
        const int x = __shfl_up(incl, o, 64);
        if (lane >= o) incl += x;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int run = incl - local;
    for (int v = 0; v < wave; ++v) run += wsum[v];
    const size_t row = static_cast<size_t>(l) * nw;
    for (int w = w0; w < w0 + per && w < nw; ++w) {
        bits[row + w] = sbits[w];
        rank[row + w] = static_cast<uint16_t>(run);
        run += __popc(sbits[w]);
    }
}

// InvertedIndex postings: the term-sorted positions become (doc, value) pairs.
__global__ void gather_postings_kernel(const int32_t* __restrict__ spos, const int32_t* __restrict__ doc_of,
                                       const float* __restrict__ values, size_t n, int32_t* __restrict__ docs,
                                       float* __restrict__ vals) {
    GRID_STRIDE(i, n) {
        const int32_t p = spos[i];
        docs[i] = doc_of[p];
        vals[i] = values[p];
    }
}

// One workgroup per term: the list's largest value from 0 up (values that
// are not > the running maximum, NaN included, leave it), and whether a doc
// repeats inside the list (a row listing the term twice) -- then `dup` is set.
__global__ __launch_bounds__(kB) void term_max_kernel(const int64_t* __restrict__ list_ptr,
                                                      const int32_t* __restrict__ docs,
                                                      const float* __restrict__ vals, float* __restrict__ max_score,
                                                      int32_t* __restrict__ dup) {
    __shared__ float wmax[kB / 64];
    const int64_t a = list_ptr[blockIdx.x], b = list_ptr[blockIdx.x + 1];
    float m = 0.0F;
    bool rep = false;
    for (int64_t i = a + threadIdx.x; i < b; i += kB) {
        const float v = vals[i];
        if (v > m) m = v;
        rep = rep || (i > a && docs[i] == docs[i - 1]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(m, off);
        if (o > m) m = o;
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    if (rep) *dup = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = wmax[0];
        for (int w = 1; w < kB / 64; ++w)
            if (wmax[w] > r) r = wmax[w];
        max_score[blockIdx.x] = r;
    }
}

__global__ void widen_i32_kernel(const int32_t* __restrict__ in, int64_t* __restrict__ out, size_t n) {
    GRID_STRIDE(i, n) out[i] = in[i];
}

}  // namespace

// ArrayInvertedLists::build_inverted_lists (invlists/inverted_lists.cpp:159-180):
// a stable radix sort of the row entries by term keeps doc order (and row
// order inside a doc) within each list; InvertedIndex::build then takes each
// list's maximum from 0 (InvertedList::max_value, inverted_lists.cpp:88-96).
void build_postings_on_device(const int32_t* indptr, const uint16_t* indices, const float* values, int64_t n_docs,
                              int64_t nnz, int dim, DevicePostings& out) {
    const size_t Z = static_cast<size_t>(nnz), L = static_cast<size_t>(dim);
    out.list_ptr.allocate((L + 1) * 8);
    out.max_score.allocate(L * 4);
    out.docs.allocate(Z * 4, 64);
    out.vals.allocate(Z * 4, 64);
    out.n_postings = nnz;
    hip_check(hipMemset(out.max_score.get(), 0, L * 4), "hipMemset");  // lists left empty keep 0
    DeviceBuffer dup(4);
    hip_check(hipMemset(dup.get(), 0, 4), "hipMemset");
    if (Z == 0) {
        hip_check(hipMemset(out.list_ptr.get(), 0, (L + 1) * 8), "hipMemset");
    } else {
        DeviceBuffer pos_iota(Z * 4), spos(Z * 4), skeys(Z * 2), hist(L * 4), hist64(L * 8), doc_of(Z * 4);
        hipLaunchKernelGGL(iota_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, pos_iota.get<int32_t>(), Z);
        radix_sort_pairs(indices, skeys.get<uint16_t>(), pos_iota.get<int32_t>(), spos.get<int32_t>(), Z, 16);
        hip_check(hipMemset(hist.get(), 0, L * 4), "hipMemset");
        hipLaunchKernelGGL(term_runs_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, skeys.get<uint16_t>(), Z,
                           hist.get<int32_t>());
        hipLaunchKernelGGL(widen_i32_kernel, dim3(grid_for(L)), dim3(kB), 0, nullptr, hist.get<int32_t>(),
                           hist64.get<int64_t>(), L);
        exclusive_scan_i64(hist64.get<int64_t>(), out.list_ptr.get<int64_t>(), L);
        hipLaunchKernelGGL(doc_of_kernel, dim3(grid_for(static_cast<size_t>(n_docs))), dim3(kB), 0, nullptr, indptr,
                           n_docs, doc_of.get<int32_t>());
        hipLaunchKernelGGL(gather_postings_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, spos.get<int32_t>(),
                           doc_of.get<int32_t>(), values, Z, out.docs.get<int32_t>(), out.vals.get<float>());
        if (L > 0)
            hipLaunchKernelGGL(term_max_kernel, dim3(static_cast<unsigned>(L)), dim3(kB), 0, nullptr,
                               out.list_ptr.get<int64_t>(), out.docs.get<int32_t>(), out.vals.get<float>(),
                               out.max_score.get<float>(), dup.get<int32_t>());
        sync("inverted postings");
    }
    int32_t h_dup = 0;
    dup.download(&h_dup, 4);
    out.duplicate_docs = h_dup != 0;
    build_window_starts(out, n_docs, dim);
}

bool build_summary_bitmaps(const DevIndexView& ix, DeviceBuffer& bits, DeviceBuffer& rank) {
    bits.release();
    rank.release();
    if (ix.n_lists <= 0 || ix.list_term_base == nullptr || ix.sum_term_ids == nullptr) return false;
    const int nw = (ix.dim + 31) / 32;
    const size_t L = static_cast<size_t>(ix.n_lists);
    // rank prefixes are 16-bit: every list must have fewer than 65536 summary terms
    std::vector<int32_t> ltb(L + 1);
    hip_check(hipMemcpy(ltb.data(), ix.list_term_base, (L + 1) * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (size_t l = 0; l < L; ++l)
        if (ltb[l + 1] - ltb[l] >= 65536) return false;
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t need = L * nw * 6;
    if (need + (size_t{1} << 30) > free_b) return false;
    bits.allocate(L * nw * 4);
    rank.allocate(L * nw * 2);
    hipLaunchKernelGGL(summary_bitmap_kernel, dim3(static_cast<unsigned>(L)), dim3(kB),
                       nw * 4 + (kB / 64) * 4, nullptr, ix.list_term_base, ix.sum_term_ids, nw,
                       bits.get<uint32_t>(), rank.get<uint16_t>());
    sync("summary bitmaps");
    return true;
}

bool build_block_rows(const DevIndexView& ix, int64_t n_postings, DeviceBuffer& row_ref, DeviceBuffer& blk_rows,
                      int64_t& n_words) {
    n_words = 0;
    row_ref.release();
    blk_rows.release();
    if (ix.esz != 1 || n_postings <= 0 || ix.n_docs <= 0) return false;
    if (ix.indptr == nullptr || ix.indices == nullptr || ix.values == nullptr || ix.docs == nullptr ||
        ix.list_cluster_base == nullptr || ix.cluster_doc_ptr == nullptr)
        throw std::logic_error("block rows need the forward index and the posting lists on the device");
    const size_t N = static_cast<size_t>(ix.n_docs), D = static_cast<size_t>(n_postings);
    const auto* values = static_cast<const uint8_t*>(ix.values);
    // each doc's row, encoded once
    DeviceBuffer dw(N * 8), dws((N + 1) * 8), flag(4), span(8);
    hip_check(hipMemset(flag.get(), 0, 4), "hipMemset");
    hip_check(hipMemset(span.get(), 0, 8), "hipMemset");
    hipLaunchKernelGGL(doc_block_words_kernel, dim3(grid_for(N)), dim3(kB), 0, nullptr, ix.indptr, ix.indices, values,
                       N, dw.get<int64_t>(), flag.get<int32_t>());
    exclusive_scan_i64(dw.get<int64_t>(), dws.get<int64_t>(), N);
    sync("block row encoding sizes");
    dw.release();
    int32_t too_long = 0;
    int64_t doc_words = 0;
    flag.download(&too_long, 4);
    dws.download(&doc_words, 8, N * 8);
    if (too_long != 0) return false;
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (static_cast<size_t>(doc_words) * 16 + D * 24 + (size_t{1} << 30) > free_b) return false;
    DeviceBuffer doc_rows(static_cast<size_t>(doc_words) * 16, 64);
    hipLaunchKernelGGL(doc_block_encode_kernel, dim3(grid_for(N)), dim3(kB), 0, nullptr, ix.indptr, ix.indices,
                       values, N, dws.get<int64_t>(), doc_rows.get<uint4>());
    // then every posting's copy of its doc's words
    DeviceBuffer words(D * 8), ws((D + 1) * 8);
    hipLaunchKernelGGL(block_words_kernel, dim3(grid_for(D)), dim3(kB), 0, nullptr, ix.docs, dws.get<int64_t>(), D,
                       words.get<int64_t>());
    exclusive_scan_i64(words.get<int64_t>(), ws.get<int64_t>(), D);
    hipLaunchKernelGGL(list_word_span_kernel, dim3(grid_for(static_cast<size_t>(ix.n_lists))), dim3(kB), 0, nullptr,
                       ix.list_cluster_base, ix.cluster_doc_ptr, ws.get<int64_t>(), ix.n_lists,
                       span.get<unsigned long long>());
    sync("block row sizes");
    words.release();
    int64_t W = 0, max_span = 0;
    ws.download(&W, 8, D * 8);
    span.download(&max_span, 8);
    // the search keeps a list's row offsets as 32-bit counts of 8ths of words
    if (max_span * 8 >= (int64_t{1} << 31)) return false;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (static_cast<size_t>(W) * 16 + D * 8 + (size_t{1} << 30) > free_b) return false;
    row_ref.allocate(D * 8);
    blk_rows.allocate(static_cast<size_t>(W) * 16, 64);
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((D + kB / 16 - 1) / (kB / 16), 1u << 20));
    hipLaunchKernelGGL(block_copy_kernel, dim3(std::max(grid, 1u)), dim3(kB), 0, nullptr, ix.indptr, ix.docs,
                       dws.get<int64_t>(), doc_rows.get<uint4>(), ws.get<int64_t>(), D, blk_rows.get<uint4>(),
                       row_ref.get<int64_t>());
    sync("block row copy");
    n_words = W;
    return true;
}

namespace {

// Splits lists [0, L) into consecutive ranges [b[i], b[i + 1]) whose entry
// counts (start[l1] - start[l0], `start` holding L + 1 prefix offsets) stay
// at or below `cap`; a single list larger than `cap` forms a range alone.
std::vector<int> list_batches(const std::vector<int64_t>& start, int64_t cap) {
    const int L = static_cast<int>(start.size()) - 1;
    std::vector<int> b{0};
    for (int l = 0; l < L; ++l)
        if (l > b.back() && start[l + 1] - start[b.back()] > cap) b.push_back(l);
    b.push_back(L);
    return b;
}

int64_t max_batch_entries(const std::vector<int64_t>& start, const std::vector<int>& b) {
    int64_t m = 0;
    for (size_t i = 0; i + 1 < b.size(); ++i) m = std::max(m, start[b[i + 1]] - start[b[i]]);
    return m;
}

// Entries per batch of the two list-keyed sorts of the build (centroid index
// and summary transpose): 32-bit sort values and a bounded scratch footprint.
// NSPARSE_BUILD_BATCH_ENTRIES lowers it (tests use it to force many batches
// on small corpora).
int64_t batch_entries() {
    constexpr int64_t kMax = int64_t{1} << 29;
    if (const char* e = std::getenv("NSPARSE_BUILD_BATCH_ENTRIES")) {
        const long long v = std::atoll(e);
        if (v > 0) return std::min<int64_t>(v, kMax);
    }
    return kMax;
}

int calc_lambda(int lambda, int64_t n) {
    if (lambda == -1) return std::max(static_cast<int>(0.0005F * static_cast<float>(n)), 160);
    return lambda;
}
int calc_beta(int beta, int lambda) {
    if (beta == -1) return static_cast<int>(static_cast<float>(lambda) * 0.1F);
    return beta;
}

}  // namespace

void encode_sq_on_device(const float* host_values, size_t n, float vmin, float vmax, uint8_t* host_codes,
                         int bits) {
    if (n == 0) return;
    if (bits != 8 && bits != 16) throw std::invalid_argument("codes are 8 or 16 bits");
    const size_t cb = bits / 8;
    const float scale = (bits == 16 ? 65535.0F : 255.0F) / (vmax - vmin);
    constexpr size_t kChunk = size_t{1} << 28;
    DeviceBuffer dv(std::min(n, kChunk) * 4), dc(std::min(n, kChunk) * cb);
    for (size_t off = 0; off < n; off += kChunk) {
        const size_t m = std::min(kChunk, n - off);
        hip_check(hipMemcpy(dv.get(), host_values + off, m * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        if (bits == 16)
            hipLaunchKernelGGL((sq_encode_kernel<uint16_t>), dim3(grid_for(m)), dim3(kB), 0, nullptr, dv.get<float>(),
                               dc.get<uint16_t>(), m, vmin, scale);
        else
            hipLaunchKernelGGL((sq_encode_kernel<uint8_t>), dim3(grid_for(m)), dim3(kB), 0, nullptr, dv.get<float>(),
                               dc.get<uint8_t>(), m, vmin, scale);
        hip_check(hipGetLastError(), "sq_encode_kernel launch");
        hip_check(hipMemcpy(host_codes + off * cb, dc.get(), m * cb, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
}

void build_on_device(const DevIndexView& fwd, int64_t n_docs, int64_t nnz, int lambda_p, int beta_p, float alpha,
                     int seed, DeviceLists& out) {
    const int L = fwd.dim;
    const bool sq = fwd.esz == 1;   // uint8 codes
    const bool w16 = fwd.esz == 2;  // uint16 codes: the FP32 flow over integer codes
    const int lambda = calc_lambda(lambda_p, n_docs);
    const int beta = calc_beta(beta_p, lambda);
    const uint32_t base_seed = seed == -1 ? std::random_device{}() : static_cast<uint32_t>(seed);
    const size_t Z = static_cast<size_t>(nnz);

    // ---- 1. postings by term (stable: doc order within a term)
    DeviceBuffer pos_iota(Z * 4), spos(Z * 4), skeys(Z * 2), hist(static_cast<size_t>(L) * 4), doc_of(Z * 4);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, pos_iota.get<int32_t>(), Z);
    radix_sort_pairs(fwd.indices, skeys.get<uint16_t>(), pos_iota.get<int32_t>(), spos.get<int32_t>(), Z, 16);
    hip_check(hipMemset(hist.get(), 0, static_cast<size_t>(L) * 4), "hipMemset");
    hipLaunchKernelGGL(term_runs_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, skeys.get<uint16_t>(), Z,
                       hist.get<int32_t>());
    hipLaunchKernelGGL(doc_of_kernel, dim3(grid_for(static_cast<size_t>(n_docs))), dim3(kB), 0, nullptr, fwd.indptr,
                       n_docs, doc_of.get<int32_t>());
    sync("postings");
    skeys.release();
    const auto h_hist = hist.to_host<int32_t>(L);
    std::vector<int64_t> list_start(L + 1, 0), kept_base(L + 1, 0);
    std::vector<int32_t> lcb(L + 1, 0);
    bool any_long = false;
    int max_clusters = 0;
    for (int l = 0; l < L; ++l) {
        const int64_t len = h_hist[l];
        list_start[l + 1] = list_start[l] + len;
        // prune_and_keep_doc_ids: lambda <= 0 (as size_t: 0 or huge) or >= len keeps all
        const int64_t kl = (lambda <= 0 || lambda >= len) ? len : lambda;
        any_long = any_long || kl < len;
        kept_base[l + 1] = kept_base[l] + kl;
        // boundary_check_n_clusters (random_kmeans.cpp:53-63), beta as size_t
        int64_t nc = 0;
        if (kl > 0) {
            if (beta == 0)
                nc = static_cast<int64_t>(std::sqrt(static_cast<double>(kl)));
            else if (beta < 0)
                nc = kl;
            else
                nc = beta;
            nc = std::min(nc, kl);
            nc = std::max<int64_t>(1, nc);
        }
        if (nc > 65536) throw std::invalid_argument("beta above 65536 clusters per list cannot be stored");
        lcb[l + 1] = lcb[l] + static_cast<int32_t>(nc);
        max_clusters = std::max<int>(max_clusters, static_cast<int>(nc));
    }
    const int64_t K = kept_base[L];
    const int64_t C = lcb[L];
    if (max_clusters > 4096)
        throw std::invalid_argument("more than 4096 clusters per list is beyond the device build's LDS plan");

    // ---- 2. lambda pruning order
    DeviceBuffer spos2;
    if (any_long) {
        spos2.allocate(Z * 4);
        if (sq) {
            DeviceBuffer k_in(Z * 4), k_out(Z * 4);
            hipLaunchKernelGGL(prune_key_u8_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, fwd.indices,
                               static_cast<const uint8_t*>(fwd.values), Z, k_in.get<uint32_t>());
            radix_sort_pairs(k_in.get<uint32_t>(), k_out.get<uint32_t>(), pos_iota.get<int32_t>(),
                             spos2.get<int32_t>(), Z, 24);
        } else if (w16) {
            DeviceBuffer k_in(Z * 4), k_out(Z * 4);
            hipLaunchKernelGGL(prune_key_u16_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, fwd.indices,
                               static_cast<const uint16_t*>(fwd.values), Z, k_in.get<uint32_t>());
            radix_sort_pairs(k_in.get<uint32_t>(), k_out.get<uint32_t>(), pos_iota.get<int32_t>(),
                             spos2.get<int32_t>(), Z, 32);
        } else {
            DeviceBuffer k_in(Z * 8), k_out(Z * 8);
            hipLaunchKernelGGL(prune_key_f32_kernel, dim3(grid_for(Z)), dim3(kB), 0, nullptr, fwd.indices,
                               static_cast<const float*>(fwd.values), Z, k_in.get<uint64_t>());
            radix_sort_pairs(k_in.get<uint64_t>(), k_out.get<uint64_t>(), pos_iota.get<int32_t>(),
                             spos2.get<int32_t>(), Z, 48);
        }
        sync("pruning sort");
    }
    pos_iota.release();
    DeviceBuffer d_list_start, d_kept_base, kept(static_cast<size_t>(K) * 4), d_lcb;
    d_list_start.upload(list_start);
    d_kept_base.upload(kept_base);
    d_lcb.upload(lcb);
    hipLaunchKernelGGL(kept_docs_kernel, dim3(std::min(L, 65536)), dim3(kB), 0, nullptr, d_list_start.get<int64_t>(),
                       d_kept_base.get<int64_t>(), spos.get<int32_t>(),
                       any_long ? spos2.get<int32_t>() : spos.get<int32_t>(), doc_of.get<int32_t>(),
                       kept.get<int32_t>(), L);
    // build counters: [0] (doc term, centroid posting) pairs scored by the
    // assignment, [1] non-zeros of the kept postings' rows
    DeviceBuffer bstat(16);
    hip_check(hipMemset(bstat.get(), 0, 16), "hipMemset");
    if (K > 0)
        hipLaunchKernelGGL(kept_row_nnz_kernel, dim3(grid_for(static_cast<size_t>(K))), dim3(kB), 0, nullptr,
                           fwd.indptr, kept.get<int32_t>(), static_cast<size_t>(K),
                           bstat.get<unsigned long long>() + 1);
    sync("kept docs");
    spos.release();
    spos2.release();
    doc_of.release();

    // ---- 3. centroid draws
    DeviceBuffer work(static_cast<size_t>(K) * 4), centroid_doc(static_cast<size_t>(C) * 4),
        cluster_list(static_cast<size_t>(C) * 4);
    hipLaunchKernelGGL(centroid_kernel, dim3(std::min(L, 65536)), dim3(64), 0, nullptr, d_kept_base.get<int64_t>(),
                       kept.get<int32_t>(), work.get<int32_t>(), d_lcb.get<int32_t>(), centroid_doc.get<int32_t>(),
                       base_seed, L);
    hipLaunchKernelGGL(cluster_list_kernel, dim3(std::min(L, 65536)), dim3(kB), 0, nullptr, d_lcb.get<int32_t>(),
                       cluster_list.get<int32_t>(), L);
    sync("centroids");
    work.release();

    // ---- 4. centroid term-major index + assignment, in batches of consecutive
    // lists so that each batch's sort values and CSC offsets fit 32 bits
    DeviceBuffer clen(static_cast<size_t>(C) * 8), cent_off(static_cast<size_t>(C + 1) * 8);
    hipLaunchKernelGGL(row_len_kernel, dim3(grid_for(C)), dim3(kB), 0, nullptr, fwd.indptr, centroid_doc.get<int32_t>(),
                       cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), d_kept_base.get<int64_t>(),
                       static_cast<size_t>(C), clen.get<int64_t>());
    exclusive_scan_i64(clen.get<int64_t>(), cent_off.get<int64_t>(), static_cast<size_t>(C));
    sync("centroid sizes");
    clen.release();
    std::vector<int64_t> list_ce(L + 1);
    {
        const auto h_cent_off = cent_off.to_host<int64_t>(static_cast<size_t>(C) + 1);
        for (int l = 0; l <= L; ++l) list_ce[l] = h_cent_off[lcb[l]];
    }
    const std::vector<int> cbatch = list_batches(list_ce, batch_entries());
    const int64_t cmax = max_batch_entries(list_ce, cbatch);
    const size_t ew = sq ? 1 : w16 ? 2 : 4;
    DeviceBuffer ckey_in(cmax * 4), cval_in(cmax * 4), ent_c(cmax * 2), ent_w(cmax * ew), ckey(cmax * 4),
        cperm(cmax * 4), cc(cmax * 2), cw(cmax * ew);
    const int assign_grid = std::min(L, 2048);
    DeviceBuffer tmap(static_cast<size_t>(assign_grid) * L * 8), assign(static_cast<size_t>(K) * 4);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(static_cast<size_t>(assign_grid) * L * 2)), dim3(kB), 0, nullptr,
                       tmap.get<int32_t>(), static_cast<size_t>(assign_grid) * L * 2, -1);
    const int mc = std::max(1, max_clusters);
    int mc2 = 1;
    while (mc2 < mc) mc2 <<= 1;
    // sims: 4 waves x mc accumulators (int64 for 16-bit codes, whose sums pass 2^32)
    const size_t assign_lds = ((mc2 * 4 + 15) & ~15) + static_cast<size_t>(4) * mc * (w16 ? 8 : 4);
    if (assign_lds > 64 * 1024) {
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(assign_kernel<uint8_t, int32_t>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(assign_lds)),
                  "hipFuncSetAttribute");
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(assign_kernel<float, float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(assign_lds)),
                  "hipFuncSetAttribute");
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(assign_kernel<uint16_t, unsigned long long>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(assign_lds)),
                  "hipFuncSetAttribute");
    }
    // uint8 lists take the hybrid assignment (dense hot terms on the matrix
    // cores) when its LDS plan keeps three workgroups per CU
    const bool hot = sq && mc <= kHotMaxClusters;
    const int hot_grid = std::min(L, 1024);
    const size_t mcp = static_cast<size_t>((mc + 15) & ~15);
    DeviceBuffer bmat;
    if (hot) {
        bmat.allocate(static_cast<size_t>(hot_grid) * mcp * kHot * 2);
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(assign_hot_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, hot_lds(mc).total),
                  "hipFuncSetAttribute");
    }
    for (size_t bi = 0; bi + 1 < cbatch.size(); ++bi) {
        const int l0 = cbatch[bi], l1 = cbatch[bi + 1];
        const size_t g0 = static_cast<size_t>(lcb[l0]), ng = static_cast<size_t>(lcb[l1]) - g0;
        const int64_t eb = list_ce[l0], n = list_ce[l1] - eb;
        if (n > 0) {
            if (sq)
                hipLaunchKernelGGL((centroid_entries_kernel<uint8_t>), dim3(grid_for(ng, kB / 16)), dim3(kB), 0, nullptr,
                                   fwd.indptr, fwd.indices, static_cast<const uint8_t*>(fwd.values),
                                   centroid_doc.get<int32_t>(), cluster_list.get<int32_t>(), d_lcb.get<int32_t>(),
                                   cent_off.get<int64_t>(), g0, ng, eb, ckey_in.get<uint32_t>(),
                                   cval_in.get<int32_t>(), ent_c.get<uint16_t>(), ent_w.get<uint8_t>());
            else if (w16)
                hipLaunchKernelGGL((centroid_entries_kernel<uint16_t>), dim3(grid_for(ng, kB / 16)), dim3(kB), 0,
                                   nullptr, fwd.indptr, fwd.indices, static_cast<const uint16_t*>(fwd.values),
                                   centroid_doc.get<int32_t>(), cluster_list.get<int32_t>(), d_lcb.get<int32_t>(),
                                   cent_off.get<int64_t>(), g0, ng, eb, ckey_in.get<uint32_t>(),
                                   cval_in.get<int32_t>(), ent_c.get<uint16_t>(), ent_w.get<uint16_t>());
            else
                hipLaunchKernelGGL((centroid_entries_kernel<float>), dim3(grid_for(ng, kB / 16)), dim3(kB), 0, nullptr,
                                   fwd.indptr, fwd.indices, static_cast<const float*>(fwd.values),
                                   centroid_doc.get<int32_t>(), cluster_list.get<int32_t>(), d_lcb.get<int32_t>(),
                                   cent_off.get<int64_t>(), g0, ng, eb, ckey_in.get<uint32_t>(),
                                   cval_in.get<int32_t>(), ent_c.get<uint16_t>(), ent_w.get<float>());
            radix_sort_pairs(ckey_in.get<uint32_t>(), ckey.get<uint32_t>(), cval_in.get<int32_t>(),
                             cperm.get<int32_t>(), static_cast<size_t>(n), 32);
            if (sq)
                hipLaunchKernelGGL((gather_csc_kernel<uint8_t>), dim3(grid_for(n)), dim3(kB), 0, nullptr,
                                   cperm.get<int32_t>(), ent_c.get<uint16_t>(), ent_w.get<uint8_t>(),
                                   static_cast<size_t>(n), cc.get<uint16_t>(), cw.get<uint8_t>());
            else if (w16)
                hipLaunchKernelGGL((gather_csc_kernel<uint16_t>), dim3(grid_for(n)), dim3(kB), 0, nullptr,
                                   cperm.get<int32_t>(), ent_c.get<uint16_t>(), ent_w.get<uint16_t>(),
                                   static_cast<size_t>(n), cc.get<uint16_t>(), cw.get<uint16_t>());
            else
                hipLaunchKernelGGL((gather_csc_kernel<float>), dim3(grid_for(n)), dim3(kB), 0, nullptr,
                                   cperm.get<int32_t>(), ent_c.get<uint16_t>(), ent_w.get<float>(),
                                   static_cast<size_t>(n), cc.get<uint16_t>(), cw.get<float>());
        }
        const int grid = std::min(l1 - l0, assign_grid);
        if (hot)
            hipLaunchKernelGGL(assign_hot_kernel, dim3(std::min(l1 - l0, hot_grid)), dim3(kHB), hot_lds(mc).total,
                               nullptr, fwd.indptr, fwd.indices, static_cast<const uint8_t*>(fwd.values),
                               d_kept_base.get<int64_t>(), kept.get<int32_t>(), d_lcb.get<int32_t>(),
                               centroid_doc.get<int32_t>(), cent_off.get<int64_t>(), eb, ckey.get<uint32_t>(),
                               cc.get<uint16_t>(), cw.get<uint8_t>(), tmap.get<int2>(), L, bmat.get<__bf16>(),
                               assign.get<int32_t>(), l0, l1, mc, bstat.get<unsigned long long>());
        else if (sq)
            hipLaunchKernelGGL((assign_kernel<uint8_t, int32_t>), dim3(grid), dim3(kB), assign_lds, nullptr,
                               fwd.indptr, fwd.indices, static_cast<const uint8_t*>(fwd.values),
                               d_kept_base.get<int64_t>(), kept.get<int32_t>(), d_lcb.get<int32_t>(),
                               centroid_doc.get<int32_t>(), cent_off.get<int64_t>(), eb, ckey.get<uint32_t>(),
                               cc.get<uint16_t>(), cw.get<uint8_t>(), tmap.get<int2>(), L, assign.get<int32_t>(), l0,
                               l1, mc, bstat.get<unsigned long long>());
        else if (w16)
            hipLaunchKernelGGL((assign_kernel<uint16_t, unsigned long long>), dim3(grid), dim3(kB), assign_lds, nullptr,
                               fwd.indptr, fwd.indices, static_cast<const uint16_t*>(fwd.values),
                               d_kept_base.get<int64_t>(), kept.get<int32_t>(), d_lcb.get<int32_t>(),
                               centroid_doc.get<int32_t>(), cent_off.get<int64_t>(), eb, ckey.get<uint32_t>(),
                               cc.get<uint16_t>(), cw.get<uint16_t>(), tmap.get<int2>(), L, assign.get<int32_t>(), l0,
                               l1, mc, bstat.get<unsigned long long>());
        else
            hipLaunchKernelGGL((assign_kernel<float, float>), dim3(grid), dim3(kB), assign_lds, nullptr, fwd.indptr,
                               fwd.indices, static_cast<const float*>(fwd.values), d_kept_base.get<int64_t>(),
                               kept.get<int32_t>(), d_lcb.get<int32_t>(), centroid_doc.get<int32_t>(),
                               cent_off.get<int64_t>(), eb, ckey.get<uint32_t>(), cc.get<uint16_t>(),
                               cw.get<float>(), tmap.get<int2>(), L, assign.get<int32_t>(), l0, l1, mc,
                               bstat.get<unsigned long long>());
        sync("assignment");
    }
    tmap.release();
    bmat.release();
    ckey_in.release();
    cval_in.release();
    ent_c.release();
    ent_w.release();
    ckey.release();
    cperm.release();
    cc.release();
    cw.release();
    cent_off.release();

    // ---- 5. clusters
    DeviceBuffer ccount(static_cast<size_t>(C) * 8), cdp(static_cast<size_t>(C + 1) * 8);
    hipLaunchKernelGGL(cluster_count_kernel, dim3(std::min(L, 65536)), dim3(kB), 0, nullptr, d_kept_base.get<int64_t>(),
                       assign.get<int32_t>(), d_lcb.get<int32_t>(), ccount.get<int64_t>(), L);
    exclusive_scan_i64(ccount.get<int64_t>(), cdp.get<int64_t>(), static_cast<size_t>(C));
    sync("cluster sizes");
    int64_t D = 0;
    cdp.download(&D, 8, static_cast<size_t>(C) * 8);
    DeviceBuffer docs(static_cast<size_t>(D) * 4);
    hipLaunchKernelGGL(cluster_place_kernel, dim3(std::min(L, 65536)), dim3(64), 0, nullptr, d_kept_base.get<int64_t>(),
                       kept.get<int32_t>(), assign.get<int32_t>(), d_lcb.get<int32_t>(), centroid_doc.get<int32_t>(),
                       cdp.get<int64_t>(), docs.get<int32_t>(), L);
    sync("cluster placement");
    kept.release();
    assign.release();
    ccount.release();
    centroid_doc.release();

    // ---- 6. summaries (count pass, scan, write pass)
    const int nw = (L + 31) / 32;
    DeviceBuffer scount(static_cast<size_t>(C) * 8), out_off(static_cast<size_t>(C + 1) * 8);
    const unsigned sgrid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(C, sq ? 8192 : 1024)));
    const size_t lds8 = static_cast<size_t>(nw) * 8 * 4 + nw * 4 + 256 * 4 + nw * 8 + 8 * 4;
    const int f32_cap = 4096;
    const size_t ldsf = static_cast<size_t>(L) * 4 + nw * 8 + 8 + f32_cap * 8 + 32;
    int spill_stride = 1;
    while (spill_stride < L) spill_stride <<= 1;
    DeviceBuffer spill;
    if (!sq) spill.allocate(static_cast<size_t>(sgrid) * spill_stride * sizeof(KeyItem));
    if (!sq && ldsf > 160 * 1024)
        throw std::invalid_argument("vocabulary too large for the FP32 / 16-bit device summarizer");
    if (!sq) {
        hip_check(hipFuncSetAttribute(w16 ? reinterpret_cast<const void*>(summarize_seq_kernel<uint16_t>)
                                          : reinterpret_cast<const void*>(summarize_seq_kernel<float>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(ldsf)),
                  "hipFuncSetAttribute");
    }
    // uint8: small clusters one wavefront each, the others (listed by the count
    // pass) one workgroup each
    DeviceBuffer big, n_big;
    const unsigned small_grid =
        static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((C + kSumWaves - 1) / kSumWaves, 8192)));
    if (sq) {
        big.allocate(static_cast<size_t>(C) * 4);
        n_big.allocate(4);
        hip_check(hipMemset(n_big.get(), 0, 4), "hipMemset");
    }
    int64_t summary_entries = 0;
    for (int pass = 0; pass < 2; ++pass) {
        DeviceBuffer sterm, sval, scl;
        int64_t E = 0;
        if (pass == 1) {
            exclusive_scan_i64(scount.get<int64_t>(), out_off.get<int64_t>(), static_cast<size_t>(C));
            sync("summary offsets");
            out_off.download(&E, 8, static_cast<size_t>(C) * 8);
            summary_entries = E;
            sterm.allocate(E * 2);
            sval.allocate(E * ew, 64);
            scl.allocate(E * 2);
        }
        if (sq) {
            const auto* vals8 = static_cast<const uint8_t*>(fwd.values);
            if (pass == 0) {
                hipLaunchKernelGGL((summarize_small_u8_kernel<false>), dim3(small_grid), dim3(kB), 0, nullptr,
                                   fwd.indptr, fwd.indices, vals8, cdp.get<int64_t>(), docs.get<int32_t>(),
                                   cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), C, alpha, scount.get<int64_t>(),
                                   nullptr, nullptr, nullptr, nullptr, big.get<int32_t>(), n_big.get<int32_t>());
                hipLaunchKernelGGL((summarize_u8_kernel<false>), dim3(sgrid), dim3(kB), lds8, nullptr, fwd.indptr,
                                   fwd.indices, vals8, cdp.get<int64_t>(), docs.get<int32_t>(),
                                   cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), C, nw, alpha,
                                   scount.get<int64_t>(), nullptr, nullptr, nullptr, nullptr, big.get<int32_t>(),
                                   n_big.get<int32_t>());
            } else {
                hipLaunchKernelGGL((summarize_small_u8_kernel<true>), dim3(small_grid), dim3(kB), 0, nullptr,
                                   fwd.indptr, fwd.indices, vals8, cdp.get<int64_t>(), docs.get<int32_t>(),
                                   cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), C, alpha, scount.get<int64_t>(),
                                   out_off.get<int64_t>(), sterm.get<uint16_t>(), sval.get<uint8_t>(),
                                   scl.get<uint16_t>(), big.get<int32_t>(), n_big.get<int32_t>());
                hipLaunchKernelGGL((summarize_u8_kernel<true>), dim3(sgrid), dim3(kB), lds8, nullptr, fwd.indptr,
                                   fwd.indices, vals8, cdp.get<int64_t>(), docs.get<int32_t>(),
                                   cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), C, nw, alpha,
                                   scount.get<int64_t>(), out_off.get<int64_t>(), sterm.get<uint16_t>(),
                                   sval.get<uint8_t>(), scl.get<uint16_t>(), big.get<int32_t>(), n_big.get<int32_t>());
            }
        } else if (w16) {
            hipLaunchKernelGGL((summarize_seq_kernel<uint16_t>), dim3(sgrid), dim3(kB), ldsf, nullptr, fwd.indptr,
                               fwd.indices, static_cast<const uint16_t*>(fwd.values), cdp.get<int64_t>(),
                               docs.get<int32_t>(), cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), C, L, f32_cap,
                               alpha, scount.get<int64_t>(), out_off.get<int64_t>(), pass == 1,
                               pass == 1 ? sterm.get<uint16_t>() : nullptr,
                               pass == 1 ? sval.get<uint16_t>() : nullptr, pass == 1 ? scl.get<uint16_t>() : nullptr,
                               spill.get<KeyItem>(), spill_stride);
        } else {
            hipLaunchKernelGGL((summarize_seq_kernel<float>), dim3(sgrid), dim3(kB), ldsf, nullptr, fwd.indptr,
                               fwd.indices, static_cast<const float*>(fwd.values), cdp.get<int64_t>(),
                               docs.get<int32_t>(), cluster_list.get<int32_t>(), d_lcb.get<int32_t>(), C, L, f32_cap,
                               alpha, scount.get<int64_t>(), out_off.get<int64_t>(), pass == 1,
                               pass == 1 ? sterm.get<uint16_t>() : nullptr, pass == 1 ? sval.get<float>() : nullptr,
                               pass == 1 ? scl.get<uint16_t>() : nullptr, spill.get<KeyItem>(), spill_stride);
        }
        sync("summaries");
        if (pass == 0) continue;

        // ---- 7. term-major transpose per list, in batches of consecutive lists
        std::vector<int64_t> list_se(L + 1);
        {
            const auto h_out_off = out_off.to_host<int64_t>(static_cast<size_t>(C) + 1);
            for (int l = 0; l <= L; ++l) list_se[l] = h_out_off[lcb[l]];
        }
        const std::vector<int> sbatch = list_batches(list_se, batch_entries());
        const int64_t smax = max_batch_entries(list_se, sbatch);
        DeviceBuffer key_in(smax * 4), val_in(smax * 4), key(smax * 4), perm(smax * 4), flag(smax * 8),
            hscan((smax + 1) * 8);
        out.sum_cluster.allocate(E * 2);
        out.sum_value.allocate(E * ew, 64);
        out.list_entry_base.allocate(static_cast<size_t>(L + 1) * 8);
        out.list_term_base.allocate(static_cast<size_t>(L + 1) * 4);
        std::vector<DeviceBuffer> head_ids, head_ptr;
        std::vector<int64_t> head_count;
        int64_t T = 0;
        for (size_t bi = 0; bi + 1 < sbatch.size(); ++bi) {
            const int l0 = sbatch[bi], l1 = sbatch[bi + 1];
            const size_t g0 = static_cast<size_t>(lcb[l0]), ng = static_cast<size_t>(lcb[l1]) - g0;
            const int64_t eb = list_se[l0], n = list_se[l1] - eb;
            if (n > 0) {
                hipLaunchKernelGGL(summary_key_kernel, dim3(grid_for(ng)), dim3(kB), 0, nullptr, sterm.get<uint16_t>(),
                                   out_off.get<int64_t>(), cluster_list.get<int32_t>(), g0, ng, eb,
                                   key_in.get<uint32_t>(), val_in.get<int32_t>());
                radix_sort_pairs(key_in.get<uint32_t>(), key.get<uint32_t>(), val_in.get<int32_t>(),
                                 perm.get<int32_t>(), static_cast<size_t>(n), 32);
                if (sq)
                    hipLaunchKernelGGL((gather_summary_kernel<uint8_t>), dim3(grid_for(n)), dim3(kB), 0, nullptr,
                                       perm.get<int32_t>(), scl.get<uint16_t>() + eb, sval.get<uint8_t>() + eb,
                                       static_cast<size_t>(n), out.sum_cluster.get<uint16_t>() + eb,
                                       out.sum_value.get<uint8_t>() + eb);
                else if (w16)
                    hipLaunchKernelGGL((gather_summary_kernel<uint16_t>), dim3(grid_for(n)), dim3(kB), 0, nullptr,
                                       perm.get<int32_t>(), scl.get<uint16_t>() + eb, sval.get<uint16_t>() + eb,
                                       static_cast<size_t>(n), out.sum_cluster.get<uint16_t>() + eb,
                                       out.sum_value.get<uint16_t>() + eb);
                else
                    hipLaunchKernelGGL((gather_summary_kernel<float>), dim3(grid_for(n)), dim3(kB), 0, nullptr,
                                       perm.get<int32_t>(), scl.get<uint16_t>() + eb, sval.get<float>() + eb,
                                       static_cast<size_t>(n), out.sum_cluster.get<uint16_t>() + eb,
                                       out.sum_value.get<float>() + eb);
                hipLaunchKernelGGL(head_flag_kernel, dim3(grid_for(n)), dim3(kB), 0, nullptr, key.get<uint32_t>(),
                                   static_cast<size_t>(n), flag.get<int64_t>());
            }
            exclusive_scan_i64(flag.get<int64_t>(), hscan.get<int64_t>(), static_cast<size_t>(n));
            sync("summary heads");
            int64_t tb = 0;
            hscan.download(&tb, 8, static_cast<size_t>(n) * 8);
            head_ids.emplace_back(tb * 2);
            head_ptr.emplace_back(tb * 8);
            head_count.push_back(tb);
            if (n > 0)
                hipLaunchKernelGGL(head_write_kernel, dim3(grid_for(n)), dim3(kB), 0, nullptr, key.get<uint32_t>(),
                                   hscan.get<int64_t>(), static_cast<size_t>(n), eb,
                                   head_ids.back().get<uint16_t>(), head_ptr.back().get<int64_t>());
            hipLaunchKernelGGL(list_bases_kernel, dim3(grid_for(l1 - l0)), dim3(kB), 0, nullptr, d_lcb.get<int32_t>(),
                               out_off.get<int64_t>(), hscan.get<int64_t>(), eb, n, T, l0, l1,
                               out.list_entry_base.get<int64_t>(), out.list_term_base.get<int32_t>());
            sync("summary transpose");
            T += tb;
            if (T >= (int64_t{1} << 31)) throw std::invalid_argument("summary term table exceeds 32-bit offsets");
        }
        out.sum_term_ids.allocate(T * 2);
        out.sum_term_ptr.allocate((T + 1) * 8);
        int64_t h = 0;
        for (size_t bi = 0; bi < head_count.size(); ++bi) {
            const int64_t tb = head_count[bi];
            if (tb > 0) {
                hip_check(hipMemcpy(out.sum_term_ids.get<uint16_t>() + h, head_ids[bi].get(), tb * 2,
                                    hipMemcpyDeviceToDevice),
                          "hipMemcpy D2D");
                hip_check(hipMemcpy(out.sum_term_ptr.get<int64_t>() + h, head_ptr[bi].get(), tb * 8,
                                    hipMemcpyDeviceToDevice),
                          "hipMemcpy D2D");
            }
            h += tb;
        }
        const int32_t t32 = static_cast<int32_t>(T);
        hip_check(hipMemcpy(out.sum_term_ptr.get<int64_t>() + T, &E, 8, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(out.list_entry_base.get<int64_t>() + L, &E, 8, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(out.list_term_base.get<int32_t>() + L, &t32, 4, hipMemcpyHostToDevice), "hipMemcpy");
        sync("summary term table");
    }
    out.list_cluster_base = std::move(d_lcb);
    out.cluster_doc_ptr = std::move(cdp);
    out.docs = std::move(docs);
    out.n_lists = L;
    out.max_clusters = max_clusters;
    uint64_t counters[2] = {0, 0};
    bstat.download(counters, 16);
    out.stats.kept_postings = K;
    out.stats.kept_row_nnz = static_cast<int64_t>(counters[1]);
    out.stats.clusters = C;
    out.stats.summary_entries = summary_entries;
    out.stats.assign_pairs = static_cast<int64_t>(counters[0]);
}

}  // namespace nsparse_amd
