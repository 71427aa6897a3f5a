"""The GPU's own forms of libstdc++'s orderings (search_common.h), run on the
GPU through the test hooks: the first-list cluster sort by one wavefront and
by a whole workgroup (one recursion level at a time over all its waves), and
the query cut with its heap held in a wavefront's registers. Each must equal
the real std::sort / std::priority_queue the oracle calls, on inputs dense
with ties (where a wrong step order changes the output)."""
import ctypes

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu


def dev_order(nsparse, keys, workgroup):
    keys = np.ascontiguousarray(keys, np.float32)
    out = np.empty(len(keys), np.int64)
    nsparse._lib.check(nsparse._lib.lib.nsparse_testing_device_cluster_order(
        keys.ctypes.data_as(ctypes.c_void_p), len(keys), workgroup, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def dev_top_k_tokens(nsparse, terms, w, k):
    terms = np.ascontiguousarray(terms, np.uint16)
    w = np.ascontiguousarray(w, np.float32)
    out = np.empty(len(terms) + 1, np.uint16)
    n = ctypes.c_int()
    nsparse._lib.check(nsparse._lib.lib.nsparse_testing_device_top_k_tokens(
        terms.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p), len(terms), k,
        out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
    return out[: n.value]


@pytest.mark.parametrize("workgroup", [0, 1])
@pytest.mark.parametrize("levels", [1, 2, 3, 8, 50, 100000])
def test_device_cluster_order_matches_std_sort(nsparse, workgroup, levels):
    rng = np.random.default_rng(levels)
    for n in list(range(0, 40)) + [63, 64, 65, 127, 200, 600, 1000, 4096]:
        keys = rng.integers(0, levels, n).astype(np.float32)
        assert (dev_order(nsparse, keys, workgroup) == O.cluster_order(keys)).all(), (levels, n)


@pytest.mark.parametrize("workgroup", [0, 1])
def test_device_cluster_order_adversarial(nsparse, workgroup):
    # sorted, reverse-sorted and organ-pipe inputs exercise the heap-sort fallback
    for n in (100, 600, 2000, 4096):
        for keys in (np.arange(n), np.arange(n)[::-1], np.r_[np.arange(n // 2), np.arange(n // 2)[::-1]],
                     np.tile([0, 1], n // 2), np.zeros(n)):
            keys = keys.astype(np.float32)
            assert (dev_order(nsparse, keys, workgroup) == O.cluster_order(keys)).all(), n


@pytest.mark.parametrize("levels", [1, 2, 4, 16, 256])
def test_device_top_k_tokens_matches_priority_queue(nsparse, levels):
    rng = np.random.default_rng(200 + levels)
    for _ in range(200):
        n = int(rng.integers(1, 513))
        k = int(rng.integers(1, 65))
        terms = rng.choice(30522, n, replace=False).astype(np.uint16)
        w = rng.integers(0, levels, n).astype(np.float32)
        assert (dev_top_k_tokens(nsparse, terms, w, k) == O.top_k_tokens_f32(terms, w, k)).all(), (n, k)
