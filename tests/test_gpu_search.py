"""GPU search parity: the HIP path (through the C ABI) against the oracle on
the same index bytes and queries. Integer (SQ) and FP32 scores must be
bit-identical; labels identical (the oracle runs in its id-tie mode, which is
the device's documented tie rule); traversal counters identical to the
oracle's except rows re-scored for docs already seen in an earlier list
(the device checks those against its top-k instead of a visited set)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

DIM = 30522


@pytest.fixture(scope="module")
def data(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=0), 20000)
    queries = synth.make_corpus(synth.query_spec(seed=1), 150)
    mixed = synth.make_corpus(synth.query_spec(seed=2, log_uniform=True), 150)
    return docs, queries, mixed


def build_pair(nsparse, docs, sq, lam=300, beta=30):
    ip, ix, v = docs
    vmax = float(v.max())
    o = O.OracleIndex(DIM, sq, lam, beta, 0.4, 0, 0.0, vmax, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    dev = nsparse.read_index_bytes(o.serialize())
    return o, dev, vmax


@pytest.fixture(scope="module")
def sq_pair(nsparse, data):
    return build_pair(nsparse, data[0], 1)


@pytest.fixture(scope="module")
def f32_pair(nsparse, data):
    return build_pair(nsparse, data[0], 0)


def params(nsparse, sq, cut, hf, vmax, sel=None, sq_range=None):
    if sq and sq_range is not None:
        p = nsparse.SeismicSQSearchParameters(sq_range[0], sq_range[1], cut, hf)
    else:
        p = nsparse.SeismicSearchParameters(cut, hf)
    if sel is not None:
        p.set_id_selector(sel)
    return p


def assert_same(got, want):
    gd, gl = got
    wd, wl = want
    assert (gl == wl).all(), np.argwhere(gl != wl)[:5]
    assert (gd == wd).all(), np.argwhere(gd != wd)[:5]


@pytest.mark.parametrize("sq", [1, 0])
@pytest.mark.parametrize("cut,hf,k", [(10, 0.8, 10), (10, 1.0, 100), (3, 0.7, 10), (1000, 0.9, 10),
                                      (10, 1e9, 10), (1, 0.5, 5)])
def test_search_matches_oracle(nsparse, data, sq_pair, f32_pair, sq, cut, hf, k):
    o, dev, vmax = sq_pair if sq else f32_pair
    for qs in (data[1], data[2]):
        ip, ix, v = qs
        want = o.search(ip, ix, v, k, cut=cut, hf=hf)
        got = dev.search(len(ip) - 1, ip, ix, v, k, params(nsparse, sq, cut, hf, vmax))
        assert_same(got, want)


@pytest.mark.parametrize("block", ["256", "512", "1024"])
@pytest.mark.parametrize("cut,hf,k", [(10, 0.8, 10), (10, 1.0, 100), (1000, 0.9, 10)])
def test_sq_every_workgroup_size_matches_oracle(nsparse, data, sq_pair, monkeypatch, block, cut, hf, k):
    """The dense-table search at 256, 512 and 1024 lanes per query (chunks of
    512, 512 and 2048 doc slots) against the oracle."""
    o, dev, vmax = sq_pair
    monkeypatch.setenv("NSPARSE_SEARCH_BLOCK", block)
    for qs in (data[1], data[2]):
        ip, ix, v = qs
        want = o.search(ip, ix, v, k, cut=cut, hf=hf)
        got = dev.search(len(ip) - 1, ip, ix, v, k, params(nsparse, 1, cut, hf, vmax))
        assert_same(got, want)


def test_sq_query_range_override(nsparse, data, sq_pair):
    o, dev, vmax = sq_pair
    ip, ix, v = data[1]
    want = o.search(ip, ix, v, 10, cut=10, hf=0.8, sq_range=(0.1, 5.0))
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, 1, 10, 0.8, vmax, sq_range=(0.1, 5.0)))
    assert_same(got, want)


@pytest.mark.parametrize("sq", [1, 0])
def test_selectors(nsparse, data, sq_pair, f32_pair, sq):
    o, dev, vmax = sq_pair if sq else f32_pair
    ip, ix, v = data[1]
    rng = np.random.default_rng(7)
    n_docs = len(data[0][0]) - 1
    allowed = np.sort(rng.choice(n_docs, n_docs // 2, replace=False)).astype(np.int32)
    banned = np.sort(rng.choice(n_docs, n_docs // 5, replace=False)).astype(np.int32)
    # enumerable selector larger than k: the normal path with a filter
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("set", allowed))
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, sq, 10, 0.9, vmax, nsparse.SetIDSelector(allowed)))
    assert_same(got, want)
    # NotIDSelector(Set): not enumerable, never exact
    inner = nsparse.SetIDSelector(banned)
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("set", banned), negate=True)
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, sq, 10, 0.9, vmax, nsparse.NotIDSelector(inner)))
    assert_same(got, want)
    # bitmap selector
    mask = np.zeros(n_docs, bool)
    mask[allowed] = True
    bm = nsparse.BitmapIDSelector(mask)
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("bitmap", bm.words, n_docs))
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, sq, 10, 0.9, vmax, bm))
    assert_same(got, want)
    # small selector (size <= k): ExactMatcher path, with a repeated id
    few = np.array([allowed[3], allowed[1], allowed[3], allowed[7]], np.int32)
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("array", few))
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, sq, 10, 0.9, vmax, nsparse.ArrayIDSelector(few)))
    assert_same(got, want)


def test_queries_with_repeated_terms_and_unsorted_rows(nsparse, sq_pair, f32_pair):
    rng = np.random.default_rng(11)
    rows = []
    for _ in range(60):
        m = int(rng.integers(1, 40))
        t = rng.choice(200, m).astype(np.uint16)  # repeats allowed, unsorted
        rows.append(t)
    ip = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.int32)
    ix = np.concatenate(rows)
    v = rng.lognormal(0, 0.7, len(ix)).astype(np.float32)
    for sq, (o, dev, vmax) in ((1, sq_pair), (0, f32_pair)):
        want = o.search(ip, ix, v, 10, cut=5, hf=0.8)
        got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, sq, 5, 0.8, vmax))
        assert_same(got, want)


def test_empty_query_and_padding(nsparse, sq_pair):
    o, dev, vmax = sq_pair
    ip = np.array([0, 0, 1], np.int32)
    ix = np.array([30000], np.uint16)  # rare term: few or no docs
    v = np.array([1.0], np.float32)
    want = o.search(ip, ix, v, 50, cut=10, hf=1.0)
    got = dev.search(2, ip, ix, v, 50, params(nsparse, 1, 10, 1.0, vmax))
    assert_same(got, want)
    assert (got[1][0] == -1).all() and (got[0][0] == -1.0).all()


@pytest.fixture(scope="module")
def big_cluster_pairs(nsparse, data):
    """lambda 3000, beta 2: clusters of ~1500 docs, each split over several
    chunks of the device traversal."""
    return {sq: build_pair(nsparse, data[0], sq, lam=3000, beta=2) for sq in (1, 0)}


@pytest.mark.parametrize("sq,block", [(1, "512"), (1, "256"), (1, "1024"), (0, "256")])
@pytest.mark.parametrize("hf,k", [(0.5, 10), (0.9, 100), (1.0, 10)])
def test_clusters_larger_than_a_chunk(nsparse, data, big_cluster_pairs, monkeypatch, sq, block, hf, k):
    """A cluster split over several chunks is tested once, at its first piece;
    when that test fails on a later list, none of its pieces is scored."""
    o, dev, vmax = big_cluster_pairs[sq]
    monkeypatch.setenv("NSPARSE_SEARCH_BLOCK", block)
    for qs in (data[1], data[2]):
        ip, ix, v = qs
        want = o.search(ip, ix, v, k, cut=10, hf=hf, with_stats=True)
        import torch

        n = len(ip) - 1
        stats = torch.zeros((n, 16), dtype=torch.int64, device="cuda")
        dd, dl = dev.search_device(torch.from_numpy(ip).cuda(), torch.from_numpy(ix.view(np.int16)).cuda(),
                                   torch.from_numpy(v).cuda(), k, params=params(nsparse, sq, 10, hf, vmax),
                                   stats=stats)
        torch.cuda.synchronize()
        assert dev.device_error() == 0
        assert_same((dd.cpu().numpy(), dl.cpu().numpy()), want[:2])
        st = stats.cpu().numpy()
        # lists, terms found, summary entries, prune tests, clusters opened, doc ids read
        assert (st[:, :6] == want[2][:, :6]).all(), np.argwhere(st[:, :6] != want[2][:, :6])[:5]


def test_longest_queries(nsparse, data, sq_pair, f32_pair):
    """Queries of 480..512 non-zeros (the per-query workgroup's maximum)
    against the oracle; 513 goes through the spread search on a uint8 index
    with block rows and is rejected on an FP32 index."""
    from nsparse import synth

    spec = synth.SynthSpec(7, DIM, 1.1, 1, 480.0, 512.0, 480, 512, 0.0, 0.7)
    ip, ix, v = synth.make_corpus(spec, 40)
    assert (np.diff(ip) <= 512).all() and np.diff(ip).max() >= 500
    for sq, (o, dev, vmax) in ((1, sq_pair), (0, f32_pair)):
        want = o.search(ip, ix, v, 10, cut=10, hf=0.8)
        got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, sq, 10, 0.8, vmax))
        assert_same(got, want)
    ip2 = np.array([0, 513], np.int32)
    ix2 = np.arange(513, dtype=np.uint16)
    v2 = np.ones(513, np.float32)
    o, dev, vmax = sq_pair
    assert_same(dev.search(1, ip2, ix2, v2, 10, params(nsparse, 1, 10, 0.8, vmax)),
                o.search(ip2, ix2, v2, 10, cut=10, hf=0.8))
    o, dev, vmax = f32_pair
    with pytest.raises(ValueError):
        dev.search(1, ip2, ix2, v2, 10, params(nsparse, 0, 10, 0.8, vmax))


def test_batched_device_mixed_k_equals_single(nsparse, data, sq_pair):
    import torch

    o, dev, vmax = sq_pair
    ip, ix, v = data[2]
    n = len(ip) - 1
    ks = np.where(np.arange(n) % 2 == 0, 10, 100).astype(np.int32)
    d_ip = torch.from_numpy(ip).cuda()
    d_ix = torch.from_numpy(ix.view(np.int16)).cuda()
    d_v = torch.from_numpy(v).cuda()
    d_ks = torch.from_numpy(ks).cuda()
    p = params(nsparse, 1, 10, 0.8, vmax)
    stats = torch.zeros((n, 16), dtype=torch.int64, device="cuda")
    dd, dl = dev.search_device(d_ip, d_ix, d_v, 100, ks=d_ks, params=p, stats=stats)
    torch.cuda.synchronize()
    assert dev.device_error() == 0
    dd, dl, st = dd.cpu().numpy(), dl.cpu().numpy(), stats.cpu().numpy()
    w10 = o.search(ip, ix, v, 10, cut=10, hf=0.8, with_stats=True)
    w100 = o.search(ip, ix, v, 100, cut=10, hf=0.8, with_stats=True)
    for q in range(n):
        wd, wl, ws = w10 if ks[q] == 10 else w100
        k = ks[q]
        assert (dl[q, :k] == wl[q]).all() and (dd[q, :k] == wd[q]).all(), q
        assert (dl[q, k:] == -1).all()
        # lists, terms found, summary entries, prune tests, clusters opened, doc ids read
        assert (st[q, :6] == ws[q, :6]).all(), (q, st[q], ws[q])
        assert st[q, 6] >= ws[q, 6]


@pytest.mark.parametrize("sq", [1, 0])
def test_idmap_search_matches_oracle(nsparse, sq):
    """IDMapIndex over a GPU-built SEISMIC index against the oracle's IDMapIndex
    (id_map_index.cpp:29-75) on the same index: labels come back external,
    selectors (Set, Array, NotIDSelector, bitmap over external ids, and small
    enumerable selectors with unknown external ids that take the exact path)
    are evaluated on external ids; scores and labels identical."""
    from nsparse import synth

    n = 3000
    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=3), n)
    vmax = float(v.max())
    ids = (np.random.default_rng(5).permutation(n).astype(np.int32) * 7 + 100)
    spec = (f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda=200|beta=20|alpha=0.4|seed=0" if sq
            else "seismic,lambda=200|beta=20|alpha=0.4|seed=0")
    im = nsparse.index_factory(DIM, "idmap," + spec)
    im.add_with_ids(n, ip, ix, v, ids)
    im.build()
    o = O.OracleIndex(DIM, sq, 200, 20, 0.4, 0, 0.0, vmax, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=4), 30)
    rng = np.random.default_rng(6)
    some = np.sort(rng.choice(ids, n // 3, replace=False)).astype(np.int32)
    banned = np.sort(rng.choice(ids, n // 5, replace=False)).astype(np.int32)
    nbits = int(ids.max()) + 1
    mask = np.zeros(nbits, bool)
    mask[some] = True
    bm = nsparse.BitmapIDSelector(mask)
    small = np.array([ids[17], ids[2999], 5, ids[400]], np.int32)  # 5 is no external id
    cases = [
        (None, None, 10),
        (nsparse.SetIDSelector(some), ("set", some), 10),
        (nsparse.ArrayIDSelector(some), ("array", some), 10),
        (nsparse.NotIDSelector(nsparse.SetIDSelector(banned)), ("set", banned), 10),
        (bm, ("bitmap", bm.words, nbits), 10),
        (nsparse.SetIDSelector(small), ("set", small), 10),   # size 4 <= k: exact path
        (nsparse.ArrayIDSelector(small), ("array", small), 3),  # size 4 > k: filtered traversal
    ]
    for dev_sel, o_sel, k in cases:
        p = nsparse.SeismicSearchParameters(10, 0.9)
        if dev_sel is not None:
            p.set_id_selector(dev_sel)
        got = im.search(30, qip, qix, qv, k, p)
        neg = dev_sel is not None and isinstance(dev_sel, nsparse.NotIDSelector)
        want = o.search_idmap(ids, qip, qix, qv, k, cut=10, hf=0.9, selector=o_sel, negate=neg)
        assert_same(got, want)


def test_sq_search_with_and_without_block_rows(nsparse, data, monkeypatch):
    """uint8 indexes score from block-contiguous row copies by default; with
    NSPARSE_BLOCK_ROWS=0 the kernel reads the forward index directly. Both must
    equal the oracle."""
    o, _, vmax = build_pair(nsparse, data[0], 1)
    blob = o.serialize()
    for flag in ("0", "1"):
        monkeypatch.setenv("NSPARSE_BLOCK_ROWS", flag)
        dev = nsparse.read_index_bytes(blob)
        for qs in (data[1], data[2]):
            ip, ix, v = qs
            want = o.search(ip, ix, v, 10, cut=10, hf=0.8)
            got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, 1, 10, 0.8, vmax))
            assert_same(got, want)


@pytest.mark.parametrize("dim", [3000, 30522, 65535])
def test_block_row_encoding_edge_cases(nsparse, dim):
    """Block rows re-encode each uint8 row into 16-byte words of 4-bit codes
    (low-term words, and based words of offsets < 4096): codes above 15 are
    split over several slots of one term, code-0 elements dropped, and a
    word starts over whenever a term leaves its base's span. Rows here are
    unsorted, codes cover 0..255 (some rows all zero), terms spread over the
    whole vocabulary and long rows reach 1500 elements. The batched search
    and the single-query (spread) search on block rows must equal the
    oracle, as must the forward-index search."""
    rng = np.random.default_rng(dim)
    n = 3000
    lens = rng.integers(1, 400, n)
    lens[::97] = 1500
    ip = np.zeros(n + 1, np.int32)
    ip[1:] = np.cumsum(lens)
    ix = np.concatenate([rng.choice(dim, size=m, replace=False) for m in lens]).astype(np.uint16)
    v = rng.uniform(0.0, 10.0, ip[-1]).astype(np.float32)
    v[ip[5]:ip[6]] = 0.0
    v[rng.random(ip[-1]) < 0.05] = 0.0
    qn = rng.integers(5, 120, 40)
    qip = np.zeros(41, np.int32)
    qip[1:] = np.cumsum(qn)
    qix = np.concatenate([rng.choice(dim, size=m, replace=False) for m in qn]).astype(np.uint16)
    qv = rng.uniform(0.01, 3.0, qip[-1]).astype(np.float32)
    o = O.OracleIndex(dim, 1, 300, 30, 0.4, 0, 0.0, 10.0, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    blob = o.serialize()
    p = nsparse.SeismicSearchParameters(10, 0.8)
    want = o.search(qip, qix, qv, 10, cut=10, hf=0.8)
    dev = nsparse.read_index_bytes(blob)
    assert_same(dev.search(40, qip, qix, qv, 10, p), want)
    for i in range(0, 40, 7):
        one = np.array([0, qip[i + 1] - qip[i]], np.int32)
        got = dev.search(1, one, qix[qip[i]:qip[i + 1]], qv[qip[i]:qip[i + 1]], 10, p)
        assert_same(got, (want[0][i:i + 1], want[1][i:i + 1]))


def test_waits_between_waves_end_with_a_one_poll_limit(nsparse, data, sq_pair, monkeypatch):
    """The waves of a search workgroup wait for each other on LDS flags, each
    wait bounded. With the bound forced to one poll the launch still ends:
    every result matches the normal search, or the launch reports error 5
    (a wait timed out); a normal search afterwards is clean."""
    import torch

    o, dev, vmax = sq_pair
    ip, ix, v = data[2]
    n = len(ip) - 1
    q = (torch.from_numpy(ip).cuda(), torch.from_numpy(ix.view(np.int16)).cuda(), torch.from_numpy(v).cuda())
    p = params(nsparse, 1, 10, 0.8, vmax)
    wd, wl = (t.cpu().numpy() for t in dev.search_device(*q, 10, params=p))
    torch.cuda.synchronize()
    assert dev.device_error() == 0
    monkeypatch.setenv("NSPARSE_SEARCH_SPIN_LIMIT", "1")
    gd, gl = (t.cpu().numpy() for t in dev.search_device(*q, 10, params=p))
    err = dev.device_error()
    assert err in (0, 5)
    if err == 0:
        assert np.array_equal(gl, wl) and np.array_equal(gd, wd)
    monkeypatch.delenv("NSPARSE_SEARCH_SPIN_LIMIT")
    gd, gl = (t.cpu().numpy() for t in dev.search_device(*q, 10, params=p))
    assert dev.device_error() == 0
    assert np.array_equal(gl, wl) and np.array_equal(gd, wd)
    assert n > 0


def test_selector_index_outside_the_bitmap_table(nsparse, data, sq_pair):
    """A per-query selector index >= the bitmap table's size is rejected by the
    kernel: error 6, that query's row padded with (-1.0f, -1), the other
    queries searched as usual (their results equal the oracle's)."""
    import torch

    o, dev, vmax = sq_pair
    ip, ix, v = data[1]
    n = len(ip) - 1
    n_docs = len(data[0][0]) - 1
    mask = np.random.default_rng(11).random(n_docs) >= 0.2
    words = np.packbits(mask.astype(np.uint8), bitorder="little")
    words = np.pad(words, (0, (-len(words)) % 4)).view(np.uint32)
    g = torch.device("cuda:0")
    bm = torch.from_numpy(words.copy()).to(g)
    table = torch.tensor([bm.data_ptr()], dtype=torch.int64, device=g)
    slots = np.where(np.arange(n) % 5 == 0, 0, -1).astype(np.int32)
    slots[3] = 1  # one past the table
    slots[7] = 1000
    q = (torch.from_numpy(ip).to(g), torch.from_numpy(ix.view(np.int16)).to(g), torch.from_numpy(v).to(g))
    p = params(nsparse, 1, 10, 0.8, vmax)
    gd, gl = dev.search_device(*q, 10, params=p, selector_of_query=torch.from_numpy(slots).to(g), bitmaps=table)
    gd, gl = gd.cpu().numpy(), gl.cpu().numpy()
    assert dev.device_error() == 6
    for r in (3, 7):
        assert (gl[r] == -1).all() and (gd[r] == -1.0).all()
    wd, wl = o.search(ip, ix, v, 10, cut=10, hf=0.8)
    sd, sl = o.search(ip, ix, v, 10, cut=10, hf=0.8, selector=("bitmap", words, n_docs))
    for r in range(n):
        if r in (3, 7):
            continue
        want = (sd[r], sl[r]) if slots[r] == 0 else (wd[r], wl[r])
        assert (gl[r] == want[1]).all() and (gd[r] == want[0]).all(), r
    # a clean launch afterwards resets the flag
    dev.search_device(*q, 10, params=p)
    torch.cuda.synchronize()
    assert dev.device_error() == 0


@pytest.mark.parametrize("name", sorted(__import__("_hand_cases").CASES))
def test_hand_derived_traversal_on_device(nsparse, name):
    """The hand-made indexes of tests/_hand_cases.py (results derived by hand
    from the reference's traversal), written by the oracle and read by the
    device: the device returns the hand-derived labels and scores (id tie
    mode)."""
    from _hand_cases import CASES, build_oracle, query_arrays

    case = CASES[name]
    dev = nsparse.read_index_bytes(build_oracle(case, 1).serialize())
    d, lab = dev.search(1, *query_arrays(case), case["k"], nsparse.SeismicSearchParameters(case["cut"], case["hf"]))
    want_l, want_d = case["want"][1]
    assert list(lab[0]) == want_l
    assert list(d[0]) == want_d


@pytest.mark.parametrize("sq", [1, 0])
def test_reference_order_index(nsparse, data, sq):
    """An index in the reference's own tie order (the oracle's libstdc++ tie
    mode, whose write_index bytes are the reference build's) read by the
    device: its results equal the oracle's id-tie search on those bytes
    exactly, and the reference path's (libstdc++ result ties) up to exact
    score ties."""
    ip, ix, v = data[0]
    vmax = float(v.max())
    o0 = O.OracleIndex(DIM, sq, 300, 30, 0.4, 0, 0.0, vmax, tie_mode=0)
    o0.add(ip, ix, v)
    o0.build()
    blob = o0.serialize()
    dev = nsparse.read_index_bytes(blob)
    o1 = O.OracleIndex.load(blob, tie_mode=1)
    for qs in (data[1], data[2]):
        got = dev.search(len(qs[0]) - 1, *qs, 10, params(nsparse, sq, 10, 0.8, vmax))
        assert_same(got, o1.search(*qs, 10, cut=10, hf=0.8))
        d0, l0 = o0.search(*qs, 10, cut=10, hf=0.8)
        for r in range(len(d0)):
            assert O.same_up_to_ties(d0[r], l0[r], got[0][r], got[1][r]), r


@pytest.mark.parametrize("sq", [1, 0])
def test_search_with_and_without_summary_bitmaps(nsparse, data, monkeypatch, sq):
    """The summary columns of a query's terms are found through per-list
    summary-term bitmaps by default, or by scanning each list's summary terms
    with NSPARSE_SUMMARY_BITMAPS=0. Both must equal the oracle, traversal
    counters included."""
    import torch

    o, _, vmax = build_pair(nsparse, data[0], sq)
    blob = o.serialize()
    for flag in ("0", "1"):
        monkeypatch.setenv("NSPARSE_SUMMARY_BITMAPS", flag)
        dev = nsparse.read_index_bytes(blob)
        for qs in (data[1], data[2]):
            ip, ix, v = qs
            want = o.search(ip, ix, v, 10, cut=10, hf=0.8, with_stats=True)
            n = len(ip) - 1
            stats = torch.zeros((n, 16), dtype=torch.int64, device="cuda")
            d, lab = dev.search_device(torch.from_numpy(ip).cuda(), torch.from_numpy(ix.view(np.int16)).cuda(),
                                       torch.from_numpy(v).cuda(), 10, params=params(nsparse, sq, 10, 0.8, vmax),
                                       stats=stats)
            assert_same((d.cpu().numpy(), lab.cpu().numpy()), want[:2])
            st = stats.cpu().numpy()
            assert (st[:, :3] == want[2][:, :3]).all()  # lists, query entries found, summary entries
