"""brutal (BrutalIndex) on the device, through the C ABI, mirroring the
reference's python_tests/test_brutal_index.py: exhaustive search reproduces the
exact ranking outright. Expected scores are the generic FP32 dot product
(utils/distance.h:29-37: products in float, summed in row order, no fused
multiply-add), so scores and labels are compared exactly; every document is a
candidate, zero scores included (brutal_index.cpp:75-96), ranked score desc
then doc id asc."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIM = 30522


def generic_scores(docs, q_idx, q_val, dim):
    """Score of every doc, summed in row order in float32."""
    ip, ix, v = docs
    dense = np.zeros(dim, np.float32)
    for t, x in zip(q_idx, q_val):  # get_dense_vector_float: the last entry of a term wins
        dense[t] = x
    n = len(ip) - 1
    lens = np.diff(ip)
    L = max(int(lens.max()), 1)
    prod = np.zeros((n, L), np.float32)
    row = np.repeat(np.arange(n), lens)
    col = np.arange(len(ix)) - np.repeat(ip[:-1], lens)
    prod[row, col] = v * dense[ix.astype(np.int64)]
    acc = np.zeros(n, np.float32)
    for j in range(L):
        acc = acc + prod[:, j]
    return acc


def expected(docs, queries, k, dim=DIM):
    qip, qix, qv = queries
    nq = len(qip) - 1
    n = len(docs[0]) - 1
    labels = np.full((nq, k), -1, np.int64)
    dists = np.full((nq, k), -1.0, np.float32)
    for q in range(nq):
        s = generic_scores(docs, qix[qip[q]:qip[q + 1]], qv[qip[q]:qip[q + 1]], dim)
        order = np.lexsort((np.arange(n), -s))[:k]
        labels[q, :len(order)] = order
        dists[q, :len(order)] = s[order]
    return labels, dists


def make(nsparse, docs, spec="brutal", ids=None):
    ip, ix, v = docs
    index = nsparse.index_factory(DIM, spec)
    if ids is None:
        index.add(len(ip) - 1, ip, ix, v)
    else:
        index.add_with_ids(len(ip) - 1, ip, ix, v, ids)
    return index


@pytest.fixture(scope="module")
def corpus(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=21), 4000)
    queries = synth.make_corpus(synth.query_spec(seed=22), 60)
    return docs, queries


def test_happy_case(nsparse, corpus):
    docs, queries = corpus
    index = make(nsparse, docs)
    assert index.num_vectors() == 4000
    for k in (10, 100):
        d, lab = index.search(len(queries[0]) - 1, *queries, k)
        wl, wd = expected(docs, queries, k)
        assert np.array_equal(lab, wl), np.argwhere(lab != wl)[:5]
        assert np.array_equal(d, wd)


def test_with_id_map(nsparse, corpus):
    docs, queries = corpus
    ids = (np.arange(4000, dtype=np.int32) * 3 + 11)
    index = make(nsparse, docs, "idmap,brutal", ids=ids)
    d, lab = index.search(len(queries[0]) - 1, *queries, 10)
    wl, wd = expected(docs, queries, 10)
    assert np.array_equal(lab, np.where(wl >= 0, ids[np.maximum(wl, 0)], -1))
    assert np.array_equal(d, wd)


def test_id_selector_is_ignored(nsparse, corpus):
    """BrutalIndex::search never reads the selector (test_brutal_index.py:64-85)."""
    docs, queries = corpus
    index = make(nsparse, docs)
    p = nsparse.SearchParameters()
    p.set_id_selector(nsparse.SetIDSelector(np.arange(50, dtype=np.int32)))
    d, lab = index.search(len(queries[0]) - 1, *queries, 10, p)
    wl, _ = expected(docs, queries, 10)
    assert np.array_equal(lab, wl)


def test_incremental_add(nsparse, corpus):
    docs, queries = corpus
    ip, ix, v = docs
    index = nsparse.index_factory(DIM, "brutal")
    for lo, hi in ((0, 1300), (1300, 2600), (2600, 4000)):
        sip = (ip[lo:hi + 1] - ip[lo]).astype(np.int32)
        index.add(hi - lo, sip, ix[ip[lo]:ip[hi]], v[ip[lo]:ip[hi]])
    assert index.num_vectors() == 4000
    d, lab = index.search(len(queries[0]) - 1, *queries, 10)
    wl, wd = expected(docs, queries, 10)
    assert np.array_equal(lab, wl) and np.array_equal(d, wd)


def test_k_larger_than_corpus(nsparse, corpus):
    docs, queries = corpus
    ip, ix, v = docs
    small = nsparse.index_factory(DIM, "brutal")
    small.add(3, ip[:4].copy(), ix[:ip[3]], v[:ip[3]])
    d, lab = small.search(len(queries[0]) - 1, *queries, 10)
    assert (lab[:, 3:] == -1).all() and (d[:, 3:] == -1.0).all()
    assert (np.sort(lab[:, :3], axis=1) == [0, 1, 2]).all()


def test_batch_matches_single_query_and_groups(nsparse, corpus, monkeypatch):
    """One batch, one query at a time, and batches cut into groups of 7 queries
    give the same bytes."""
    docs, queries = corpus
    qip, qix, qv = queries
    index = make(nsparse, docs)
    bd, bl = index.search(len(qip) - 1, *queries, 10)
    for q in range(0, len(qip) - 1, 9):
        sd, sl = index.search(1, np.array([0, qip[q + 1] - qip[q]], np.int32), qix[qip[q]:qip[q + 1]],
                              qv[qip[q]:qip[q + 1]], 10)
        assert np.array_equal(sd[0], bd[q]) and np.array_equal(sl[0], bl[q])
    monkeypatch.setenv("NSPARSE_BRUTE_GROUP_QUERIES", "7")
    gd, gl = index.search(len(qip) - 1, *queries, 10)
    assert np.array_equal(gd, bd) and np.array_equal(gl, bl)


def test_zero_scores_ties_and_ascending_scores(nsparse):
    """A query sharing no term with any doc gets the first k docs at score 0;
    equal scores rank by doc id; scores rising with the doc id (every round of
    the top-k filter full of candidates) still give the last k docs."""
    n = 20000
    ip = np.arange(n + 1, dtype=np.int32)
    ix = np.zeros(n, np.uint16)
    rising = make(nsparse, (ip, ix, (np.arange(n) + 1).astype(np.float32)))
    flat = make(nsparse, (ip, ix, np.ones(n, np.float32)))
    q = (np.array([0, 1, 2], np.int32), np.array([0, 5], np.uint16), np.array([1.0, 2.0], np.float32))
    for k in (1, 10, 1000):
        d, lab = rising.search(2, *q, k)
        assert (lab[0] == np.arange(n - 1, n - 1 - k, -1)).all()
        assert (d[0] == np.arange(n, n - k, -1).astype(np.float32)).all()
        assert (lab[1] == np.arange(k)).all() and (d[1] == 0.0).all()
        d, lab = flat.search(2, *q, k)
        assert (lab[0] == np.arange(k)).all() and (d[0] == 1.0).all()
    with pytest.raises(ValueError):
        flat.search(2, *q, 1025)


def test_empty_index_leaves_outputs(nsparse):
    """Without vectors the per-query results are empty and Index::search copies
    nothing (index.cpp:38-48)."""
    index = nsparse.index_factory(DIM, "brutal")
    q = (np.array([0, 1], np.int32), np.array([3], np.uint16), np.array([1.0], np.float32))
    d, lab = index.search(1, *q, 5)
    assert d.shape == (1, 5) and lab.shape == (1, 5)
