// Differentiable exact re-scoring of mined candidates (training-style use of
// the SEISMIC scorer, compute_similarity in seismic_common.h:78-103 on FP32).
//
//   score[q, c] = sum_e d_e * Q[term(e)]     (e over the entries of row cand[q, c])
//   dL/dqv[i]   = sum_c g[q, c] * d(cand[q, c])[t_i]   if i is the last entry of term t_i, else 0
//   dL/dd[e]   += g[q, c] * Q[term(e)]
//
// Q is the query as a dense vector: a term repeated in the query takes the
// value of its last entry, as SparseVectors::get_dense_vector builds it for
// compute_similarity and as the SEISMIC and brute-force searches score it, so
// the earlier entries of a repeated term have no influence and get gradient 0.
//
// One workgroup per query. The query's distinct terms sit in LDS (bitmap +
// popcount rank table) with the value of each term's last entry and a
// per-term gradient accumulator; one wavefront per candidate row gathers the
// row from HBM once for the forward and once for the backward. Gradients of
// document weights are scatter-added with float atomics (a document can be a
// candidate of several queries), those of query weights through LDS.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

#include "device_mem.h"
#include "kernels.h"

namespace nsparse_amd {

namespace {
constexpr int kRB = 256;
constexpr int kRQMax = 512;

struct RescoreLds {
    uint32_t* qbits;
    uint16_t* qpref;
    float* qval;    // per distinct term: the value of its last entry
    float* gsum;    // per distinct term: accumulated gradient
    int32_t* qlast; // per distinct term: index of its last entry
    int32_t* wsum;
};

__device__ __forceinline__ int rank_of(const RescoreLds& L, uint32_t t, bool& present) {
    const uint32_t w = L.qbits[t >> 5];
    const uint32_t bit = 1u << (t & 31);
    present = (w & bit) != 0;
    return L.qpref[t >> 5] + __popc(w & (bit - 1));
}

// Terms >= dim are ignored (they match no document term). Callers keep qn <= kRQMax.
__device__ void setup_query(RescoreLds& L, unsigned char* smem, int nw, const uint16_t* qi, const float* qv, int qn) {
    L.qbits = reinterpret_cast<uint32_t*>(smem);
    L.qpref = reinterpret_cast<uint16_t*>(L.qbits + nw);
    L.qval = reinterpret_cast<float*>(smem + ((nw * 6 + 15) & ~15));
    L.gsum = L.qval + kRQMax;
    L.qlast = reinterpret_cast<int32_t*>(L.gsum + kRQMax);
    L.wsum = L.qlast + kRQMax;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int w = tid; w < nw; w += kRB) L.qbits[w] = 0;
    for (int r = tid; r < kRQMax; r += kRB) {
        L.qval[r] = 0.0F;
        L.gsum[r] = 0.0F;
        L.qlast[r] = -1;
    }
    __syncthreads();
    for (int i = tid; i < qn; i += kRB)
        if ((qi[i] >> 5) < nw) atomicOr(&L.qbits[qi[i] >> 5], 1u << (qi[i] & 31));
    __syncthreads();
    const int per = (nw + kRB - 1) / kRB;
    const int w0 = tid * per;
    int local = 0;
    for (int w = w0; w < w0 + per && w < nw; ++w) local += __popc(L.qbits[w]);
    int incl = local;
    for (int o = 1; o < 64; o <<= 1) {
        const int x = __shfl_up(incl, o, 64);
        if (lane >= o) incl += x;
    }
    if (lane == 63) L.wsum[wave] = incl;
    __syncthreads();
    int run = incl - local;
    for (int w = 0; w < wave; ++w) run += L.wsum[w];
    for (int w = w0; w < w0 + per && w < nw; ++w) {
        L.qpref[w] = static_cast<uint16_t>(run);
        run += __popc(L.qbits[w]);
    }
    __syncthreads();
    for (int i = tid; i < qn; i += kRB) {
        if ((qi[i] >> 5) >= nw) continue;
        bool p;
        atomicMax(&L.qlast[rank_of(L, qi[i], p)], i);
    }
    __syncthreads();
    for (int r = tid; r < kRQMax; r += kRB)
        if (L.qlast[r] >= 0) L.qval[r] = qv[L.qlast[r]];
    __syncthreads();
}

__global__ __launch_bounds__(kRB) void rescore_fwd_kernel(const int32_t* __restrict__ q_indptr,
                                                          const uint16_t* __restrict__ q_indices,
                                                          const float* __restrict__ q_values, int n_cand,
                                                          const int32_t* __restrict__ cand,
                                                          const int32_t* __restrict__ d_indptr,
                                                          const uint16_t* __restrict__ d_indices,
                                                          const float* __restrict__ d_values, int nw,
                                                          int64_t n_docs, float* __restrict__ scores,
                                                          int32_t* __restrict__ err) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t q = blockIdx.x;
    const int32_t s0 = q_indptr[q], qn = q_indptr[q + 1] - s0;
    if (qn > kRQMax || qn < 0) {
        if (threadIdx.x == 0) atomicExch(err, 1);
        return;
    }
    RescoreLds L;
    setup_query(L, smem, nw, q_indices + s0, q_values + s0, qn);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < n_cand; c += kRB / 64) {
        const int32_t d = cand[q * n_cand + c];
        float acc = 0.0F;
        if (d >= 0 && d < n_docs) {
            for (int32_t p = d_indptr[d] + lane; p < d_indptr[d + 1]; p += 64) {
                bool pr;
                const uint32_t t = d_indices[p];
                if ((t >> 5) >= static_cast<uint32_t>(nw)) continue;
                const int r = rank_of(L, t, pr);
                if (pr) acc += d_values[p] * L.qval[r];
            }
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) scores[q * n_cand + c] = acc;
    }
}

__global__ __launch_bounds__(kRB) void rescore_bwd_kernel(const int32_t* __restrict__ q_indptr,
                                                          const uint16_t* __restrict__ q_indices,
                                                          const float* __restrict__ q_values, int n_cand,
                                                          const int32_t* __restrict__ cand,
                                                          const int32_t* __restrict__ d_indptr,
                                                          const uint16_t* __restrict__ d_indices,
                                                          const float* __restrict__ d_values, int nw,
                                                          int64_t n_docs, int32_t* __restrict__ err,
                                                          const float* __restrict__ grad_scores,
                                                          float* __restrict__ grad_q, float* __restrict__ grad_d) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t q = blockIdx.x;
    const int32_t s0 = q_indptr[q], qn = q_indptr[q + 1] - s0;
    if (qn > kRQMax || qn < 0) {
        if (threadIdx.x == 0) atomicExch(err, 1);
        return;
    }
    RescoreLds L;
    setup_query(L, smem, nw, q_indices + s0, q_values + s0, qn);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < n_cand; c += kRB / 64) {
        const int32_t d = cand[q * n_cand + c];
        if (d < 0 || d >= n_docs) continue;
        const float g = grad_scores[q * n_cand + c];
        if (grad_d == nullptr && grad_q == nullptr) continue;
        for (int32_t p = d_indptr[d] + lane; p < d_indptr[d + 1]; p += 64) {
            bool pr;
            const uint32_t t = d_indices[p];
            if ((t >> 5) >= static_cast<uint32_t>(nw)) continue;
            const int r = rank_of(L, t, pr);
            if (!pr) continue;
            if (grad_d != nullptr) atomicAdd(&grad_d[p], g * L.qval[r]);
            if (grad_q != nullptr) atomicAdd(&L.gsum[r], g * d_values[p]);
        }
    }
    __syncthreads();
    if (grad_q != nullptr)
        for (int i = threadIdx.x; i < qn; i += kRB) {
            bool p;
            const uint32_t t = q_indices[s0 + i];
            float gi = 0.0F;
            if ((t >> 5) < static_cast<uint32_t>(nw)) {
                const int r = rank_of(L, t, p);
                gi = L.qlast[r] == i ? L.gsum[r] : 0.0F;
            }
            grad_q[s0 + i] = gi;
        }
}

int lds_bytes(int nw) { return ((nw * 6 + 15) & ~15) + kRQMax * 12 + 4 * 4; }

}  // namespace

void launch_rescore(bool backward, int64_t n_q, const int32_t* q_indptr, const uint16_t* q_indices,
                    const float* q_values, int n_cand, const int32_t* cand, int64_t n_docs, const int32_t* d_indptr,
                    const uint16_t* d_indices, const float* d_values, int dim, float* scores,
                    const float* grad_scores, float* grad_q, float* grad_d, int32_t* err, hipStream_t stream) {
    const int nw = (dim + 31) / 32;
    const dim3 grid(static_cast<unsigned>(n_q));
    if (!backward)
        hipLaunchKernelGGL(rescore_fwd_kernel, grid, dim3(kRB), lds_bytes(nw), stream, q_indptr, q_indices, q_values,
                           n_cand, cand, d_indptr, d_indices, d_values, nw, n_docs, scores, err);
    else
        hipLaunchKernelGGL(rescore_bwd_kernel, grid, dim3(kRB), lds_bytes(nw), stream, q_indptr, q_indices, q_values,
                           n_cand, cand, d_indptr, d_indices, d_values, nw, n_docs, err, grad_scores, grad_q, grad_d);
    hip_check(hipGetLastError(), "rescore kernel launch");
}

}  // namespace nsparse_amd
