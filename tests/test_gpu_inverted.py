"""inverted (InvertedIndex) on the device, through the C ABI, mirroring the
reference's tests/inverted_index_test.cpp and python_tests/test_inverted_index.py.

The device scores every posting of the query's terms, term by term in query
order, so a doc's score is the float32 sum of its products in that order; the
expected results (TermMajor in _oracle.py) sum the same way in numpy, and
scores and labels are compared exactly. Against the oracle's MaxScore
traversal (the reference's summation order) and float64 brute force, scores
agree within 1e-5 relative and labels differ only at such near-ties. The
device build's file bytes equal the oracle's, whose lists are pinned to the
reference's own list-building code."""
import numpy as np
import pytest

import _oracle as O
from _oracle import OracleInverted, TermMajor

pytestmark = pytest.mark.gpu

DIM = 30522


def make(nsparse, docs, spec="inverted", ids=None, dim=DIM):
    ip, ix, v = docs
    index = nsparse.index_factory(dim, spec)
    if ids is None:
        index.add(len(ip) - 1, ip, ix, v)
    else:
        index.add_with_ids(len(ip) - 1, ip, ix, v, ids)
    index.build()
    return index


def search(index, queries, k, params=None):
    ip, ix, v = queries
    return index.search(len(ip) - 1, ip, ix, v, k, params)


def maps_corpus(rows):
    ip, ix, v = [0], [], []
    for r in rows:
        for t in sorted(r):
            ix.append(t)
            v.append(r[t])
        ip.append(len(ix))
    return np.array(ip, np.int32), np.array(ix, np.uint16), np.array(v, np.float32)


def one(terms, weights):
    return np.array([0, len(terms)], np.int32), np.array(terms, np.uint16), np.array(weights, np.float32)


@pytest.fixture(scope="module")
def corpus(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=31), 30000)  # 4 windows of 8192 doc ids
    queries = synth.make_corpus(synth.query_spec(seed=32), 50)
    return docs, queries


@pytest.fixture(scope="module")
def built(nsparse, corpus):
    return make(nsparse, corpus[0])


@pytest.fixture(scope="module")
def major(corpus):
    return TermMajor(corpus[0], DIM)


@pytest.mark.parametrize("k", [1, 10, 100, 1024])
def test_exact_in_query_order(nsparse, corpus, built, major, k):
    print(f"inverted 30000 docs x 50 queries, k={k}", flush=True)
    d, lab = search(built, corpus[1], k)
    wl, wd = major.top_k(corpus[1], k)
    assert np.array_equal(lab, wl), np.argwhere(lab != wl)[:5]
    assert np.array_equal(d, wd)


def test_matches_oracle_and_brute_force(nsparse, corpus, built):
    docs, queries = corpus
    o = OracleInverted(DIM)
    o.add(*docs)
    o.build()
    od, ol = o.search(*queries, 10)
    d, lab = search(built, queries, 10)
    np.testing.assert_allclose(d, od, rtol=1e-5, atol=1e-6)
    O.assert_topk_matches(lab, d, O.exact_scores64(docs, queries, DIM), 10)
    O.assert_topk_matches(ol, od, O.exact_scores64(docs, queries, DIM), 10)


def test_file_bytes_equal_oracle_and_read_back(nsparse, corpus, built):
    docs, queries = corpus
    o = OracleInverted(DIM)
    o.add(*docs)
    o.build()
    raw = nsparse.serialize_index(built)
    assert raw == o.serialize()
    back = nsparse.read_index_bytes(raw)
    assert back.num_vectors() == 30000
    d0, l0 = search(built, queries, 10)
    d1, l1 = search(back, queries, 10)
    assert np.array_equal(l0, l1) and np.array_equal(d0, d1)


def test_mmap_file_round_trip(nsparse, corpus, built, tmp_path):
    path = str(tmp_path / "inverted.idx")
    nsparse.write_index(built, path)
    back = nsparse.read_index(path, nsparse.kUseMmap)
    d0, l0 = search(built, corpus[1], 10)
    d1, l1 = search(back, corpus[1], 10)
    assert np.array_equal(l0, l1) and np.array_equal(d0, d1)


def test_selectors(nsparse, corpus, built, major):
    docs, queries = corpus
    rng = np.random.default_rng(33)
    allowed = rng.random(30000) < 0.25
    ids = np.nonzero(allowed)[0].astype(np.int32)
    for sel, mask in [(nsparse.SetIDSelector(ids), allowed), (nsparse.ArrayIDSelector(ids), allowed),
                      (nsparse.NotIDSelector(nsparse.SetIDSelector(ids)), ~allowed),
                      (nsparse.BitmapIDSelector(allowed), allowed)]:
        p = nsparse.SearchParameters()
        p.set_id_selector(sel)
        d, lab = search(built, queries, 10, p)
        wl, wd = major.top_k(queries, 10, allowed=mask)
        assert np.array_equal(lab, wl) and np.array_equal(d, wd)
    # a SEISMIC parameter class carries the selector just the same
    p = nsparse.SeismicSearchParameters(10, 0.8)
    p.set_id_selector(nsparse.SetIDSelector(ids[:3]))
    _, lab = search(built, queries, 10, p)
    assert np.isin(lab[lab >= 0], ids[:3]).all()


def test_with_id_map(nsparse, corpus, major):
    docs, queries = corpus
    ext = (np.arange(30000, dtype=np.int32) * 7 + 1000)[::-1].copy()
    index = make(nsparse, docs, "idmap,inverted", ids=ext)
    d, lab = search(index, queries, 10)
    wl, wd = major.top_k(queries, 10)
    assert np.array_equal(lab, np.where(wl >= 0, ext[wl], -1)) and np.array_equal(d, wd)
    keep = ext[::3]
    p = nsparse.SearchParameters()
    p.set_id_selector(nsparse.SetIDSelector(keep))
    d, lab = search(index, queries, 10, p)
    allowed = np.zeros(30000, bool)
    allowed[::3] = True
    wl, wd = major.top_k(queries, 10, allowed=allowed)
    assert np.array_equal(lab, np.where(wl >= 0, ext[wl], -1)) and np.array_equal(d, wd)


def test_reference_known_answers(nsparse):
    multi = [{0: 1.0, 1: 0.5}, {0: 0.8, 2: 0.6}, {1: 0.9, 3: 0.7}]
    index = make(nsparse, maps_corpus(multi), dim=4)
    d, lab = search(index, one([0, 1], [1.0, 0.8]), 3)
    assert list(lab[0]) == [0, 1, 2]
    np.testing.assert_allclose(d[0], [1.4, 0.8, 0.72], rtol=4 * 2.0 ** -23)
    p = nsparse.SearchParameters()
    p.set_id_selector(nsparse.SetIDSelector(np.array([1, 2], np.int32)))
    d, lab = search(index, one([0, 1], [1.0, 0.8]), 3, p)
    assert list(lab[0]) == [1, 2, -1]
    index = make(nsparse, maps_corpus([{0: 1.0}, {0: 0.5}]), dim=3)
    d, lab = search(index, one([0], [1.0]), 5)
    assert list(lab[0]) == [0, 1, -1, -1, -1] and list(d[0][2:]) == [-1.0] * 3
    d, lab = search(index, one([2], [1.0]), 1)
    assert lab[0, 0] == -1
    index = make(nsparse, maps_corpus([{0: 0.3}, {0: 1.0}, {0: 0.5}]), dim=3)
    p = nsparse.SearchParameters()
    p.set_id_selector(nsparse.NotIDSelector(nsparse.SetIDSelector(np.array([1], np.int32))))
    d, lab = search(index, one([0], [1.0]), 3, p)
    assert list(lab[0]) == [2, 0, -1]
    np.testing.assert_array_equal(d[0][:2], np.float32([0.5, 0.3]))


def test_multi_window_regression(nsparse):
    """inverted_index_test.cpp:429-477 and :633-678: 6000 docs, two terms."""
    rows = []
    for i in range(6000):
        r = {0: 0.1 + (i % 10) * 0.01}
        if i % 3 == 0:
            r[1] = 0.5 + (i % 7) * 0.05
        rows.append(r)
    docs = maps_corpus(rows)
    index = make(nsparse, docs, dim=2)
    q = one([0, 1], [1.0, 1.0])
    d, lab = search(index, q, 10)
    wl, wd = TermMajor(docs, 2).top_k(q, 10)
    assert np.array_equal(lab, wl) and np.array_equal(d, wd)
    allowed = [5, 4095, 4096, 4099, 5999]
    p = nsparse.SearchParameters()
    p.set_id_selector(nsparse.SetIDSelector(np.array(allowed, np.int32)))
    _, lab = search(index, q, 10, p)
    assert sorted(lab[0][lab[0] >= 0]) == allowed


def test_edges(nsparse):
    """Empty rows and queries, a query repeating a term, negative weights (only
    positive scores are results), k beyond the corpus, several adds before one
    build, and a build consuming only what was added since the last one."""
    rng = np.random.default_rng(34)
    rows = []
    for i in range(9000):
        m = int(rng.integers(0, 12))
        t = rng.choice(300, size=m, replace=False)
        rows.append({int(a): float(b) for a, b in zip(t, rng.normal(0.5, 1.0, m))})
    docs = maps_corpus(rows)
    index = nsparse.index_factory(300, "inverted")
    ip, ix, v = docs
    for lo, hi in [(0, 4000), (4000, 4001), (4001, 9000)]:
        index.add(hi - lo, ip[lo:hi + 1] - ip[lo], ix[ip[lo]:ip[hi]], v[ip[lo]:ip[hi]])
    index.build()
    assert index.num_vectors() == 9000
    qip = np.array([0, 0, 3, 8], np.int32)
    qix = np.array([5, 5, 17, 1, 2, 3, 4, 250], np.uint16)
    qv = np.array([0.7, 0.3, -1.0, 1.0, 0.5, -0.25, 2.0, 1.5], np.float32)
    queries = (qip, qix, qv)
    d, lab = search(index, queries, 20)
    wl, wd = TermMajor(docs, 300).top_k(queries, 20)
    assert np.array_equal(lab, wl) and np.array_equal(d, wd)
    assert (lab[0] == -1).all()
    small = make(nsparse, maps_corpus([{1: 1.0}, {}, {1: 2.0}]), dim=300)
    d, lab = search(small, queries, 6)
    assert list(lab[2][:2]) == [2, 0] and (lab[2][2:] == -1).all()
    small.add(1, np.array([0, 1], np.int32), np.array([1], np.uint16), np.array([9.0], np.float32))
    small.build()  # the new lists hold only the doc added since the last build, as doc 0
    assert small.num_vectors() == 4
    d, lab = search(small, queries, 3)
    assert list(lab[2]) == [0, -1, -1] and d[2][0] == np.float32(9.0)


def test_rows_repeating_a_term(nsparse):
    """A row listing a term twice puts the doc twice into that list; both
    postings count (the reference's window scoring adds both)."""
    ip = np.array([0, 3, 5, 6], np.int32)
    ix = np.array([4, 4, 9, 2, 2, 4], np.uint16)
    v = np.array([1.0, 2.0, 0.5, 0.25, 0.25, 0.125], np.float32)
    index = make(nsparse, (ip, ix, v), dim=16)
    d, lab = search(index, one([4, 2], [1.0, 2.0]), 3)
    assert list(lab[0]) == [0, 1, 2]
    np.testing.assert_allclose(d[0], [3.0, 1.0, 0.125], rtol=1e-6)


def test_larger_corpus(nsparse):
    from nsparse import synth

    print("inverted 200000 docs x 20 queries", flush=True)
    docs = synth.make_corpus(synth.doc_spec(seed=35), 200000)
    queries = synth.make_corpus(synth.query_spec(seed=36, log_uniform=True), 20)
    index = make(nsparse, docs)
    d, lab = search(index, queries, 100)
    wl, wd = TermMajor(docs, DIM).top_k(queries, 100)
    assert np.array_equal(lab, wl) and np.array_equal(d, wd)


def test_slabs_of_several_windows(nsparse, corpus, built, major):
    """1100 queries over 4 windows: each workgroup walks a slab of 2 windows,
    carrying its cursors from one window to the next."""
    from nsparse import synth

    queries = synth.make_corpus(synth.query_spec(seed=37), 1100)
    d, lab = search(built, queries, 10)
    wl, wd = major.top_k(queries, 10)
    assert np.array_equal(lab, wl) and np.array_equal(d, wd)


def test_class_constructors_build_and_search(nsparse, corpus, major):
    """IDMapIndex(InvertedIndex(dim)) and SeismicIndex(dim) with the class
    defaults (lambda and beta derived from the corpus) build and search."""
    docs, queries = corpus
    ext = np.arange(30000, dtype=np.int32) + 500
    index = nsparse.IDMapIndex(nsparse.InvertedIndex(DIM))
    index.add_with_ids(len(docs[0]) - 1, *docs, ext)
    index.build()
    d, lab = search(index, queries, 10)
    wl, wd = major.top_k(queries, 10)
    assert np.array_equal(lab, np.where(wl >= 0, ext[wl], -1)) and np.array_equal(d, wd)
    s = nsparse.SeismicIndex(DIM, nsparse.SeismicClusterParameters(seed=0))
    s.add(len(docs[0]) - 1, *docs)
    s.build()
    o = O.OracleIndex(DIM, 0, -1, -1, 0.4, 0, tie_mode=1)
    o.add(*docs)
    o.build()
    assert nsparse.serialize_index(s) == o.serialize()
