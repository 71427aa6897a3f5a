"""Deterministic synthetic corpora (nsparse_synth_* in the C ABI).

Presets follow the workloads this index is measured on: vocabulary 30522,
Zipf(1.1) term ids without repeats inside a row, FP32 lognormal(0, 0.7)
weights; documents have lognormal lengths (median 120, sigma 0.35, clipped to
8..600); queries lognormal (median 40, sigma 0.64, clipped to 5..300) or
log-uniform over 5..300.
"""
import ctypes

import numpy as np

from ._lib import SynthSpec, check, lib

VOCAB = 30522


def doc_spec(seed=0, dim=VOCAB):
    return SynthSpec(seed, dim, 1.1, 0, 120.0, 0.35, 8, 600, 0.0, 0.7)


def query_spec(seed=1, dim=VOCAB, log_uniform=False):
    if log_uniform:
        return SynthSpec(seed, dim, 1.1, 1, 5.0, 300.0, 5, 300, 0.0, 0.7)
    return SynthSpec(seed, dim, 1.1, 0, 40.0, 0.64, 5, 300, 0.0, 0.7)


def make_corpus(spec, n_rows, first_row=0, threads=0):
    """Returns (indptr int32, indices uint16, values float32) for rows
    [first_row, first_row + n_rows) of the corpus `spec` describes."""
    lens = np.empty(n_rows, dtype=np.int32)
    check(lib.nsparse_synth_row_nnz(ctypes.byref(spec), first_row, n_rows, lens.ctypes.data_as(ctypes.c_void_p),
                                    threads))
    indptr64 = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(lens, out=indptr64[1:])
    nnz = int(indptr64[-1])
    indices = np.empty(nnz, dtype=np.uint16)
    values = np.empty(nnz, dtype=np.float32)
    check(lib.nsparse_synth_fill(ctypes.byref(spec), first_row, n_rows, indptr64.ctypes.data_as(ctypes.c_void_p),
                                 indices.ctypes.data_as(ctypes.c_void_p), values.ctypes.data_as(ctypes.c_void_p),
                                 threads))
    if nnz >= 2**31:
        raise ValueError("corpus too large for int32 offsets")
    return indptr64.astype(np.int32), indices, values
