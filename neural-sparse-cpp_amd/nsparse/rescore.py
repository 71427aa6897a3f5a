"""Differentiable exact re-scoring of SEISMIC candidates (torch.autograd).

``SparseRescore.apply(q_values, d_values, q_indptr, q_indices, candidates,
d_indptr, d_indices, dim)`` returns scores[q, c] = <query q, doc candidates[q, c]>
for FP32 sparse queries and documents held on the GPU. Its backward is a
hand-written HIP kernel: d score / d q_t = d_t and d score / d d_t = q_t,
scatter-added into gradients shaped like the two value arrays.
"""
import ctypes

import torch

from ._lib import check, lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class SparseRescore(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q_values, d_values, q_indptr, q_indices, candidates, d_indptr, d_indices, dim):
        n_q, n_cand = candidates.shape
        q_values = q_values.contiguous()
        d_values = d_values.contiguous()
        scores = torch.empty((n_q, n_cand), dtype=torch.float32, device=q_values.device)
        err = torch.zeros(1, dtype=torch.int32, device=q_values.device)
        check(lib.nsparse_rescore_device(0, n_q, q_indptr.data_ptr(), q_indices.data_ptr(), q_values.data_ptr(),
                                         n_cand, candidates.data_ptr(), d_indptr.numel() - 1, d_indptr.data_ptr(),
                                         d_indices.data_ptr(), d_values.data_ptr(), int(dim), scores.data_ptr(),
                                         None, None, None, err.data_ptr(), _stream()))
        if int(err.item()) != 0:
            raise ValueError("a query has more than 512 non-zeros")
        ctx.save_for_backward(q_values, d_values, q_indptr, q_indices, candidates, d_indptr, d_indices)
        ctx.dim = int(dim)
        return scores

    @staticmethod
    def backward(ctx, grad_scores):
        q_values, d_values, q_indptr, q_indices, candidates, d_indptr, d_indices = ctx.saved_tensors
        n_q, n_cand = candidates.shape
        g = grad_scores.contiguous().float()
        gq = torch.empty_like(q_values) if ctx.needs_input_grad[0] else None
        gd = torch.zeros_like(d_values) if ctx.needs_input_grad[1] else None
        err = torch.zeros(1, dtype=torch.int32, device=q_values.device)
        check(lib.nsparse_rescore_device(1, n_q, q_indptr.data_ptr(), q_indices.data_ptr(), q_values.data_ptr(),
                                         n_cand, candidates.data_ptr(), d_indptr.numel() - 1, d_indptr.data_ptr(),
                                         d_indices.data_ptr(), d_values.data_ptr(), ctx.dim, None, g.data_ptr(),
                                         gq.data_ptr() if gq is not None else None,
                                         gd.data_ptr() if gd is not None else None, err.data_ptr(), _stream()))
        return gq, gd, None, None, None, None, None, None


def rescore(q_values, d_values, q_indptr, q_indices, candidates, d_indptr, d_indices, dim):
    return SparseRescore.apply(q_values, d_values, q_indptr, q_indices, candidates, d_indptr, d_indices, dim)


def info_nce(scores, valid, temperature=0.05):
    """InfoNCE over candidates, the first candidate of each query positive."""
    logits = (scores / temperature).masked_fill(~valid, float("-inf"))
    return torch.nn.functional.cross_entropy(logits, torch.zeros(len(scores), dtype=torch.long,
                                                                  device=scores.device))
