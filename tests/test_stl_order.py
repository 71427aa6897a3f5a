"""The library's step-for-step re-implementation of libstdc++'s std::sort and
std::priority_queue (stl_order.h, used on the GPU for the first-list cluster
order and the query cut) against the real libstdc++ routines the oracle calls,
on inputs dense with ties."""
import ctypes

import numpy as np
import pytest

import _oracle as O


def lib_cluster_order(nsparse, keys):
    keys = np.ascontiguousarray(keys, np.float32)
    out = np.empty(len(keys), np.int64)
    nsparse._lib.check(nsparse._lib.lib.nsparse_testing_cluster_order(
        keys.ctypes.data_as(ctypes.c_void_p), len(keys), out.ctypes.data_as(ctypes.c_void_p)))
    return out


def lib_top_k_tokens(nsparse, terms, w, k):
    terms = np.ascontiguousarray(terms, np.uint16)
    w = np.ascontiguousarray(w, np.float32)
    out = np.empty(len(terms) + 1, np.uint16)
    n = ctypes.c_int()
    nsparse._lib.check(nsparse._lib.lib.nsparse_testing_top_k_tokens(
        terms.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p), len(terms), k,
        out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
    return out[: n.value]


@pytest.mark.parametrize("levels", [1, 2, 3, 8, 50, 100000])
def test_cluster_order_matches_std_sort(nsparse, levels):
    rng = np.random.default_rng(levels)
    for n in list(range(0, 40)) + [63, 64, 65, 127, 200, 600, 1000, 4096]:
        keys = rng.integers(0, levels, n).astype(np.float32)
        assert (lib_cluster_order(nsparse, keys) == O.cluster_order(keys)).all(), (levels, n)


def test_cluster_order_adversarial(nsparse):
    # sorted, reverse-sorted and organ-pipe inputs exercise the heap-sort fallback
    for n in (100, 600, 2000):
        for keys in (np.arange(n), np.arange(n)[::-1], np.r_[np.arange(n // 2), np.arange(n // 2)[::-1]],
                     np.tile([0, 1], n // 2), np.zeros(n)):
            keys = keys.astype(np.float32)
            assert (lib_cluster_order(nsparse, keys) == O.cluster_order(keys)).all()


@pytest.mark.parametrize("levels", [1, 2, 4, 16, 256])
def test_top_k_tokens_matches_priority_queue(nsparse, levels):
    rng = np.random.default_rng(100 + levels)
    for _ in range(300):
        n = int(rng.integers(1, 300))
        k = int(rng.integers(1, 40))
        terms = rng.choice(30522, n, replace=False).astype(np.uint16)
        w = rng.integers(0, levels, n).astype(np.float32)
        assert (lib_top_k_tokens(nsparse, terms, w, k) == O.top_k_tokens_f32(terms, w, k)).all()
