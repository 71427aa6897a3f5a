/*
 * Hooks that expose the library's re-implementations of libstdc++ orderings,
 * on the host and on the GPU, so tests can compare them with std::ranges::sort
 * and std::priority_queue, and a device read-bandwidth probe for measurement
 * (no reference counterpart: test and measurement surface only).
 */
#ifndef NSPARSE_TESTING_H
#define NSPARSE_TESTING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cluster visiting order of the first list (seismic_common.h:105-115):
 * writes the indices 0..n-1 sorted by keys descending, ties as libstdc++ leaves them. */
int nsparse_testing_cluster_order(const float* keys, int n, int64_t* out_order);

/* top_k_tokens (utils/vector_process.h:18-30) on float weights: writes the cut
 * terms best-first and their count. */
int nsparse_testing_top_k_tokens(const uint16_t* terms, const float* weights, int n, int k, uint16_t* out,
                                 int* out_n);

/* The first-list order as the GPU search computes it, on the GPU (host arrays,
 * n <= 4096): workgroup 0 runs the one-wavefront sort, 1 the whole-workgroup one. */
int nsparse_testing_device_cluster_order(const float* keys, int n, int workgroup, int64_t* out_order);

/* The query cut as the GPU search computes it (one wavefront, registers), on
 * the GPU: n <= 512, 1 <= k <= 64. */
int nsparse_testing_device_top_k_tokens(const uint16_t* terms, const float* weights, int n, int k, uint16_t* out,
                                        int* out_n);

/* One streaming read of a DEVICE buffer (bytes a multiple of 16, 16-byte
 * aligned) on `stream` (hipStream_t, NULL = default), without synchronizing:
 * *d_out (device uint32) receives the XOR of its 32-bit words. Times the
 * device's plain read rate, the ceiling an HBM-bound kernel is compared with. */
int nsparse_testing_read_probe(const void* d_buf, size_t bytes, uint32_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NSPARSE_TESTING_H */
