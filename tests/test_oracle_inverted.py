"""Pins the oracle's InvertedIndex (oracle/oracle.cpp, inv_*) before the device
path is compared with it.

  * build + writer: the posting lists and per-term maxima equal the reference's
    own ArrayInvertedLists::build_inverted_lists / InvertedList::max_value
    compiled into oracle/_ref, byte for byte in the INVT layout
    (test_lists_match_reference).
  * search: the hand-computed cases of the reference's
    tests/inverted_index_test.cpp, and the reference python tests' contract
    (python_tests/test_inverted_index.py): the MaxScore traversal returns the
    exact top-k of positive scores, ranked score desc then id asc, within the
    allowed set when a selector is given -- checked against float64 brute
    force over corpora spanning several 4096-doc windows, for k = 1, 10, 100.
The traversal's summation order (sorted bounds, essential then non-essential
terms) is the reference's by construction; with no compiled InvertedIndex here
(its sources need Abseil) the float bits of a score beyond the brute-force
tolerance are "parity unpinned".
"""
import numpy as np
import pytest

import _oracle as O
from _oracle import OracleInverted

needs_ref = pytest.mark.skipif(O.ref is None, reason="reference build-side library not built here")


def docs_from_maps(rows):
    ip, ix, v = [0], [], []
    for r in rows:
        for t in sorted(r):
            ix.append(t)
            v.append(r[t])
        ip.append(len(ix))
    return np.array(ip, np.int32), np.array(ix, np.uint16), np.array(v, np.float32)


def built(dim, rows, tie_mode=1):
    o = OracleInverted(dim, tie_mode)
    o.add(*docs_from_maps(rows))
    o.build()
    return o


def q(terms, weights):
    return np.array([0, len(terms)], np.int32), np.array(terms, np.uint16), np.array(weights, np.float32)


def corpus(n, dim, nnz, seed, zipf=1.1, negative=False):
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, dim + 1) ** zipf
    p /= p.sum()
    ip, ix, v = [0], [], []
    for _ in range(n):
        m = int(np.clip(round(np.exp(np.log(nnz) + 0.35 * rng.standard_normal())), 1, dim))
        t = np.sort(rng.choice(dim, size=m, replace=False, p=p))
        w = rng.lognormal(0, 0.7, m).astype(np.float32)
        if negative:
            w[rng.random(m) < 0.2] *= -1
        ix.append(t)
        v.append(w)
        ip.append(ip[-1] + m)
    return np.array(ip, np.int32), np.concatenate(ix).astype(np.uint16), np.concatenate(v).astype(np.float32)


MULTI = [{0: 1.0, 1: 0.5}, {0: 0.8, 2: 0.6}, {1: 0.9, 3: 0.7}]


# ---- inverted_index_test.cpp known answers ----

def test_search_returns_nothing_when_not_built():
    o = OracleInverted(5)
    d, lab = o.search(*q([0, 1], [1.0, 0.5]), 5, fill=(7.0, 7))
    assert (lab == 7).all() and (d == 7.0).all()


def test_add_counts_empty_vectors():
    o = built(5, [{0: 1.0}, {}, {2: 0.8}, {}])
    assert o.num_vectors() == 4


def test_finds_matching_doc_and_sorts_by_score():
    o = built(3, [{0: 1.0}])
    d, lab = o.search(*q([0], [1.0]), 1)
    assert lab[0, 0] == 0 and d[0, 0] == np.float32(1.0)
    o = built(3, [{0: 0.5}, {0: 1.0}, {0: 0.7}])
    _, lab = o.search(*q([0], [1.0]), 3)
    assert list(lab[0]) == [1, 2, 0]


def test_multi_term_dot_product():
    o = built(4, MULTI)
    d, lab = o.search(*q([0, 1], [1.0, 0.8]), 3)
    assert list(lab[0]) == [0, 1, 2]
    np.testing.assert_allclose(d[0], [1.4, 0.8, 0.72], rtol=4 * 2.0 ** -23)


def test_no_matching_term_and_k_larger_than_docs():
    o = built(3, [{0: 1.0}, {0: 0.5}])
    d, lab = o.search(*q([2], [1.0]), 1)
    assert lab[0, 0] == -1 and d[0, 0] == -1.0
    d, lab = o.search(*q([0], [1.0]), 5)
    assert list(lab[0]) == [0, 1, -1, -1, -1]
    assert list(d[0][2:]) == [-1.0] * 3


def test_selectors():
    o = built(3, [{0: 0.3}, {0: 1.0}, {0: 0.5}])
    d, lab = o.search(*q([0], [1.0]), 3, selector=("set", [0, 2]))
    assert list(lab[0]) == [2, 0, -1]
    np.testing.assert_array_equal(d[0][:2], np.float32([0.5, 0.3]))
    _, lab = o.search(*q([0], [1.0]), 3, selector=("set", [0]))
    assert list(lab[0]) == [0, -1, -1]
    _, lab = o.search(*q([0], [1.0]), 2, selector=("set", [7, 8]))
    assert list(lab[0]) == [-1, -1]
    _, lab = o.search(*q([0], [1.0]), 3, selector=("set", [1]), negate=True)
    assert list(lab[0]) == [2, 0, -1]
    o = built(4, MULTI)
    d, lab = o.search(*q([0, 1], [1.0, 0.8]), 3, selector=("set", [1, 2]))
    assert list(lab[0]) == [1, 2, -1]
    np.testing.assert_allclose(d[0][:2], [0.8, 0.72], rtol=4 * 2.0 ** -23)


def window_rows():
    rows = []
    for i in range(6000):
        r = {0: 0.1 + (i % 10) * 0.01}
        if i % 3 == 0:
            r[1] = 0.5 + (i % 7) * 0.05
        rows.append(r)
    return rows


def test_multi_window():
    o = built(2, window_rows())
    d, lab = o.search(*q([0, 1], [1.0, 1.0]), 10)
    assert lab[0, 0] >= 0 and d[0, 0] > 0
    assert np.all(np.diff(d[0]) <= 0)
    allowed = [5, 4095, 4096, 4099, 5999]
    _, lab = o.search(*q([0, 1], [1.0, 1.0]), 10, selector=("set", allowed))
    got = lab[0][lab[0] >= 0]
    assert sorted(got) == sorted(allowed)


# ---- the reference's layout, from its own list-building code ----

@needs_ref
@pytest.mark.parametrize("n,dim,nnz,seed,negative", [(300, 64, 8, 1, False), (5000, 512, 20, 2, False),
                                                     (800, 128, 6, 3, True)])
def test_lists_match_reference(n, dim, nnz, seed, negative):
    ip, ix, v = corpus(n, dim, nnz, seed, negative=negative)
    v[::17] = 0.0  # zero postings stay in the lists; maxima start from 0
    o = OracleInverted(dim)
    o.add(ip, ix, v)
    o.build()
    assert o.serialize() == O.ref_inverted_bytes(dim, ip, ix, v)


def test_unbuilt_layout():
    o = OracleInverted(9)
    o.add(np.array([0, 1], np.int32), np.array([3], np.uint16), np.array([1.0], np.float32))
    b = o.serialize()
    # fourcc, dim, num_vectors, n_terms = 0, element width 4, no maxima
    assert b == b"INVT" + np.int32(9).tobytes() + np.array([1, 0, 4, 0], np.uint64).tobytes()


# ---- the exact top-k contract, against float64 brute force ----

@pytest.mark.parametrize("k", [1, 10, 100])
def test_equals_brute_force(k):
    dim = 2000
    docs = corpus(9000, dim, 30, seed=5)
    queries = corpus(40, dim, 12, seed=6, zipf=0.9)
    o = OracleInverted(dim)
    o.add(*docs)
    o.build()
    d, lab = o.search(*queries, k)
    O.assert_topk_matches(lab, d, O.exact_scores64(docs, queries, dim), k)


def test_equals_brute_force_within_selector():
    dim = 2000
    docs = corpus(9000, dim, 30, seed=7)
    queries = corpus(30, dim, 12, seed=8, zipf=0.9)
    o = OracleInverted(dim)
    o.add(*docs)
    o.build()
    rng = np.random.default_rng(9)
    allowed = rng.random(9000) < 0.3
    ids = np.nonzero(allowed)[0].astype(np.int32)
    s64 = O.exact_scores64(docs, queries, dim)
    d, lab = o.search(*queries, 10, selector=("set", ids))
    O.assert_topk_matches(lab, d, s64, 10, allowed=allowed)
    d, lab = o.search(*queries, 10, selector=("set", ids), negate=True)
    O.assert_topk_matches(lab, d, s64, 10, allowed=~allowed)


def test_incremental_add_matches_single_batch():
    dim = 500
    docs = corpus(3000, dim, 15, seed=10)
    queries = corpus(20, dim, 8, seed=11)
    a = OracleInverted(dim)
    a.add(*docs)
    a.build()
    b = OracleInverted(dim)
    ip, ix, v = docs
    for lo, hi in [(0, 1000), (1000, 2000), (2000, 3000)]:
        b.add(ip[lo:hi + 1] - ip[lo], ix[ip[lo]:ip[hi]], v[ip[lo]:ip[hi]])
    b.build()
    assert b.num_vectors() == 3000
    assert a.serialize() == b.serialize()
    np.testing.assert_array_equal(a.search(*queries, 10)[1], b.search(*queries, 10)[1])


def test_build_rejects_out_of_range_terms():
    o = OracleInverted(4)
    o.add(np.array([0, 1], np.int32), np.array([4], np.uint16), np.array([1.0], np.float32))
    with pytest.raises(RuntimeError, match="term_id out of range"):
        o.build()
