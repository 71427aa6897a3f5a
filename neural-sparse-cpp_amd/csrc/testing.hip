// Test surface for the GPU emulations of libstdc++'s orderings
// (search_common.h): they run here on host arrays, so tests can compare them
// with the real std::sort and std::priority_queue the oracle calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>

#include "device_mem.h"
#include "kernels.h"
#include "search_common.h"
#include "stl_order.h"

namespace nsparse_amd {

namespace {

constexpr int kOrderBlock = 1024;

__host__ __device__ int order_cap(int n) { return 2 * n / 17 + 2; }

// LDS: A (KeyItem n) | Lpos, Rasc (uint16 n each) | order (uint16 n) |
// leaf_bits | range lists (int4 2 * cap) | counters
int order_lds(int n) {
    const int a = n * 8, pos = ((n * 6) + 15) & ~15, leaf = (((n + 64) / 32) * 4 + 15) & ~15;
    return a + pos + leaf + 2 * order_cap(n) * 16 + 16;
}

template <bool BLOCK>
__global__ __launch_bounds__(kOrderBlock) void cluster_order_kernel(const float* keys, int n, int64_t* out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    KeyItem* A = reinterpret_cast<KeyItem*>(smem);
    uint16_t* Lpos = reinterpret_cast<uint16_t*>(A + n);
    uint16_t* Rasc = Lpos + n;
    uint16_t* order = Rasc + n;
    uint32_t* leaf = reinterpret_cast<uint32_t*>(smem + n * 8 + (((n * 6) + 15) & ~15));
    int4* rng = reinterpret_cast<int4*>(reinterpret_cast<unsigned char*>(leaf) + ((((n + 64) / 32) * 4 + 15) & ~15));
    int32_t* cnt = reinterpret_cast<int32_t*>(rng + 2 * order_cap(n));
    for (int i = tid; i < n; i += kOrderBlock) A[i] = KeyItem{keys[i], static_cast<uint32_t>(i)};
    __syncthreads();
    if constexpr (BLOCK) {
        introsort_desc_block<kOrderBlock>(A, n, Lpos, Rasc, leaf, rng, order_cap(n), cnt);
    } else {
        if (tid < 64) introsort_desc_wave(A, n, Lpos, Rasc, leaf, tid);
        __syncthreads();
    }
    introsort_final_pass(A, n, leaf, order, tid, kOrderBlock);
    __syncthreads();
    for (int i = tid; i < n; i += kOrderBlock) out[i] = order[i];
}

__global__ __launch_bounds__(64) void top_k_tokens_kernel(const uint16_t* terms, const float* weights, int n, int k,
                                                          uint16_t* out, int32_t* out_n) {
    __shared__ float w[kQMax];
    __shared__ uint16_t t[kQMax];
    __shared__ uint16_t o[kQMax];
    const int lane = threadIdx.x;
    for (int i = lane; i < n; i += 64) {
        w[i] = weights[i];
        t[i] = terms[i];
    }
    __syncthreads();
    const int cnt = topk_tokens_wave(n, k, w, t, o, lane);
    __syncthreads();
    for (int i = lane; i < cnt; i += 64) out[i] = o[i];
    if (lane == 0) *out_n = cnt;
}

}  // namespace

void device_cluster_order(const float* keys, int n, int64_t* out_order, bool block) {
    if (n < 0 || n > kMaxTestOrder) throw std::invalid_argument("device cluster order takes at most 4096 keys");
    if (n == 0) return;
    DeviceBuffer dk(static_cast<size_t>(n) * 4), dout(static_cast<size_t>(n) * 8);
    hip_check(hipMemcpy(dk.get(), keys, static_cast<size_t>(n) * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
    const int lds = order_lds(n);
    const void* fn = block ? reinterpret_cast<const void*>(cluster_order_kernel<true>)
                           : reinterpret_cast<const void*>(cluster_order_kernel<false>);
    hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds), "hipFuncSetAttribute");
    if (block)
        hipLaunchKernelGGL(cluster_order_kernel<true>, dim3(1), dim3(kOrderBlock), lds, nullptr, dk.get<float>(), n,
                           dout.get<int64_t>());
    else
        hipLaunchKernelGGL(cluster_order_kernel<false>, dim3(1), dim3(kOrderBlock), lds, nullptr, dk.get<float>(), n,
                           dout.get<int64_t>());
    hip_check(hipGetLastError(), "cluster order launch");
    dout.download(out_order, static_cast<size_t>(n) * 8);
}

namespace {

// Streams a buffer once and XORs its words into *out: the device's streaming
// read rate, for the bandwidth a search's HBM traffic is compared with. Of the
// shapes measured on MI355X (16-byte loads two to eight in flight per lane,
// grid-stride or a contiguous chunk per workgroup, 8 to 64 workgroups per CU,
// plain or non-temporal loads), four non-temporal loads in flight with 64
// workgroups per CU read fastest.
typedef unsigned int probe_u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void read_probe_kernel(const probe_u32x4* __restrict__ p, size_t n16,
                                                         uint32_t* __restrict__ out) {
    __shared__ uint32_t part[4];
    uint32_t acc = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const probe_u32x4 a = __builtin_nontemporal_load(p + i);
        const probe_u32x4 b = __builtin_nontemporal_load(p + i + stride);
        const probe_u32x4 c = __builtin_nontemporal_load(p + i + 2 * stride);
        const probe_u32x4 d = __builtin_nontemporal_load(p + i + 3 * stride);
        const probe_u32x4 x = a ^ b ^ c ^ d;
        acc ^= x.x ^ x.y ^ x.z ^ x.w;
    }
    for (; i < n16; i += stride) {
        const probe_u32x4 a = __builtin_nontemporal_load(p + i);
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    for (int o = 32; o > 0; o >>= 1) acc ^= __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicXor(out, part[0] ^ part[1] ^ part[2] ^ part[3]);
}

__global__ void zero_word_kernel(uint32_t* p) { *p = 0; }

}  // namespace

void device_read_probe(const void* d_buf, size_t bytes, uint32_t* d_out, hipStream_t stream) {
    if (d_buf == nullptr || d_out == nullptr) throw std::invalid_argument("unexpected nullptr");
    if (bytes % 16 != 0 || reinterpret_cast<uintptr_t>(d_buf) % 16 != 0)
        throw std::invalid_argument("the read probe takes a 16-byte aligned buffer of whole 16-byte words");
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    hipLaunchKernelGGL(zero_word_kernel, dim3(1), dim3(1), 0, stream, d_out);
    hipLaunchKernelGGL(read_probe_kernel, dim3(static_cast<unsigned>(std::max(cus, 1) * 64)), dim3(256), 0, stream,
                       static_cast<const probe_u32x4*>(d_buf), bytes / 16, d_out);
    hip_check(hipGetLastError(), "read probe launch");
}

int device_top_k_tokens(const uint16_t* terms, const float* weights, int n, int k, uint16_t* out) {
    if (n < 0 || n > kQMax || k <= 0 || k > 64)
        throw std::invalid_argument("device top_k_tokens takes at most 512 entries and k in [1, 64]");
    DeviceBuffer dt(static_cast<size_t>(std::max(n, 1)) * 2), dw(static_cast<size_t>(std::max(n, 1)) * 4),
        dout(static_cast<size_t>(std::max(n, 1)) * 2), dn(4);
    if (n > 0) {
        hip_check(hipMemcpy(dt.get(), terms, static_cast<size_t>(n) * 2, hipMemcpyHostToDevice), "hipMemcpy H2D");
        hip_check(hipMemcpy(dw.get(), weights, static_cast<size_t>(n) * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    hipLaunchKernelGGL(top_k_tokens_kernel, dim3(1), dim3(64), 0, nullptr, dt.get<uint16_t>(), dw.get<float>(), n, k,
                       dout.get<uint16_t>(), dn.get<int32_t>());
    hip_check(hipGetLastError(), "top_k_tokens launch");
    int32_t cnt = 0;
    dn.download(&cnt, 4);
    dout.download(out, static_cast<size_t>(cnt) * 2);
    return cnt;
}

}  // namespace nsparse_amd
