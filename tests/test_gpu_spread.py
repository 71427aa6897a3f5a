"""GPU parity of the spread search (search_spread.hip): each query of a small
launch on the whole GPU -- one workgroup per cut list for its summary scores
and first-list order, several per list for its rows' scores -- then an exact
replay of the reference's walk. It runs for launches of at most
NSPARSE_SPREAD_MAX_QUERIES queries (default 32) over uint8 indexes with block
rows, so the plain drop-in (one query per call) takes it. These tests run
queries one per call and also force it for whole batches; every result must
equal the oracle's (scores bit for bit, labels in the id tie mode), as the
per-query workgroup's do."""
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


def build_pair(nsparse, docs, lam=300, beta=30):
    ip, ix, v = docs
    vmax = float(v.max())
    o = O.OracleIndex(DIM, 1, lam, beta, 0.4, 0, 0.0, vmax, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    dev = nsparse.read_index_bytes(o.serialize())
    return o, dev, vmax


@pytest.fixture(scope="module")
def pair(nsparse, data):
    return build_pair(nsparse, data[0])


@pytest.fixture
def forced(monkeypatch):
    """The spread search for launches of any size."""
    monkeypatch.setenv("NSPARSE_SPREAD_MAX_QUERIES", "1000000")


def params(nsparse, cut, hf, sel=None):
    p = nsparse.SeismicSearchParameters(cut, hf)
    if sel is not None:
        p.set_id_selector(sel)
    return p


def one_by_one(dev, qs, k, p, n=None):
    """The plain drop-in: one query per call."""
    ip, ix, v = qs
    n = len(ip) - 1 if n is None else n
    D = np.empty((n, k), np.float32)
    L = np.empty((n, k), np.int32)
    for i in range(n):
        a, b = int(ip[i]), int(ip[i + 1])
        d, lab = dev.search(1, np.array([0, b - a], np.int32), ix[a:b], v[a:b], k, p)
        D[i], L[i] = d[0], lab[0]
    return D, L


def assert_same(got, want):
    gd, gl = got
    wd, wl = want
    assert (gl == wl).all(), np.argwhere(gl != wl)[:5]
    assert (gd == wd).all(), np.argwhere(gd != wd)[:5]


@pytest.mark.parametrize("cut,hf,k", [(10, 0.8, 10), (10, 1.0, 100), (3, 0.7, 10), (64, 0.9, 10), (10, 1e9, 10),
                                      (1, 0.5, 5), (10, 0.3, 1)])
def test_one_query_per_call_matches_oracle(nsparse, data, pair, cut, hf, k):
    o, dev, _ = pair
    for qs in (data[1], data[2]):
        want = o.search(*qs, k, cut=cut, hf=hf)
        got = one_by_one(dev, qs, k, params(nsparse, cut, hf), n=50)
        assert_same(got, (want[0][:50], want[1][:50]))


# k of 1, 2, 63 and 65 around the 64-key staging of the replay's merges (every
# staged key forces an exact test at k = 1; more than 64 keys span several
# merges), and k of 200 and 500 with tight and loose pruning
@pytest.mark.parametrize("cut,hf,k", [(10, 0.8, 10), (10, 1.0, 100), (3, 0.7, 10), (64, 0.9, 10), (1, 0.5, 5),
                                      (10, 0.5, 1), (10, 0.6, 2), (10, 0.7, 63), (10, 0.9, 65), (10, 0.4, 200),
                                      (10, 0.95, 500)])
def test_forced_batches_match_oracle(nsparse, data, pair, forced, cut, hf, k):
    o, dev, _ = pair
    for qs in (data[1], data[2]):
        want = o.search(*qs, k, cut=cut, hf=hf)
        got = dev.search(len(qs[0]) - 1, *qs, k, params(nsparse, cut, hf))
        assert_same(got, want)


def test_selectors(nsparse, data, pair, forced):
    o, dev, _ = pair
    ip, ix, v = data[1]
    rng = np.random.default_rng(7)
    n_docs = len(data[0][0]) - 1
    allowed = np.sort(rng.choice(n_docs, n_docs // 2, replace=False)).astype(np.int32)
    banned = np.sort(rng.choice(n_docs, n_docs // 5, replace=False)).astype(np.int32)
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("set", allowed))
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, 10, 0.9, nsparse.SetIDSelector(allowed)))
    assert_same(got, want)
    inner = nsparse.SetIDSelector(banned)
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("set", banned), negate=True)
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, 10, 0.9, nsparse.NotIDSelector(inner)))
    assert_same(got, want)
    mask = np.zeros(n_docs, bool)
    mask[allowed] = True
    bm = nsparse.BitmapIDSelector(mask)
    want = o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("bitmap", bm.words, n_docs))
    got = dev.search(len(ip) - 1, ip, ix, v, 10, params(nsparse, 10, 0.9, bm))
    assert_same(got, want)


def test_repeated_terms_unsorted_rows_and_empty_queries(nsparse, pair, forced):
    o, dev, _ = pair
    rng = np.random.default_rng(11)
    rows = [rng.choice(200, int(rng.integers(1, 40))).astype(np.uint16) for _ in range(60)]
    rows[5] = np.zeros(0, np.uint16)
    rows[6] = np.array([30000], np.uint16)  # a rare term: few or no docs
    ip = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.int32)
    ix = np.concatenate(rows)
    v = rng.lognormal(0, 0.7, len(ix)).astype(np.float32)
    for k in (10, 50):
        want = o.search(ip, ix, v, k, cut=5, hf=0.8)
        got = dev.search(len(ip) - 1, ip, ix, v, k, params(nsparse, 5, 0.8))
        assert_same(got, want)
    assert (got[1][5] == -1).all() and (got[0][5] == -1.0).all()


def test_queries_longer_than_the_workgroup_tables(nsparse, data, pair):
    """Queries of more than 512 non-zeros (the per-query workgroup's tables)
    go through the spread search, whose tables are sized for the call's
    longest query: single calls and a batch mixing short and long queries
    equal the oracle; a query longer than its LDS can hold is rejected."""
    o, dev, _ = pair
    rng = np.random.default_rng(23)
    lens = [700, 1500, 12, 2400, 513]
    rows = [np.sort(rng.choice(DIM, m, replace=False)).astype(np.uint16) for m in lens]
    ip = np.r_[0, np.cumsum(lens)].astype(np.int32)
    ix = np.concatenate(rows)
    v = rng.lognormal(0, 0.7, len(ix)).astype(np.float32)
    for k in (10, 100):
        want = o.search(ip, ix, v, k, cut=10, hf=0.8)
        got = dev.search(len(lens), ip, ix, v, k, params(nsparse, 10, 0.8))
        assert_same(got, want)
        for i in (0, 1, 3):
            one = np.array([0, lens[i]], np.int32)
            got1 = dev.search(1, one, rows[i], v[ip[i]:ip[i + 1]], k, params(nsparse, 10, 0.8))
            assert_same(got1, (want[0][i:i + 1], want[1][i:i + 1]))
    huge = np.sort(rng.choice(DIM, 20000, replace=False)).astype(np.uint16)
    with pytest.raises(ValueError):
        dev.search(1, np.array([0, len(huge)], np.int32), huge, np.ones(len(huge), np.float32), 10,
                   params(nsparse, 10, 0.8))


@pytest.mark.parametrize("lam,beta", [(6000, 2), (20000, 10)])
def test_large_clusters_and_long_lists(nsparse, data, forced, lam, beta):
    """Clusters of thousands of docs (a list workgroup scores its slice in
    several passes) and lists longer than the replay stages in LDS (their doc
    keys are read from HBM)."""
    o, dev, _ = build_pair(nsparse, data[0], lam=lam, beta=beta)
    for qs in (data[1], data[2]):
        for hf, k in ((0.5, 10), (0.9, 100)):
            want = o.search(*qs, k, cut=10, hf=hf)
            got = dev.search(len(qs[0]) - 1, *qs, k, params(nsparse, 10, hf))
            assert_same(got, want)


def test_device_mixed_k_and_per_query_bitmaps(nsparse, data, pair, forced):
    """search_device with per-query k and per-query bitmap selectors, one
    selector index past the table: that query's row is padded and error 6 is
    reported, every other row equals the oracle's."""
    import torch

    o, dev, _ = pair
    ip, ix, v = data[2]
    n = len(ip) - 1
    n_docs = len(data[0][0]) - 1
    mask = np.random.default_rng(11).random(n_docs) >= 0.2
    words = np.packbits(mask.astype(np.uint8), bitorder="little")
    words = np.pad(words, (0, (-len(words)) % 4)).view(np.uint32)
    g = torch.device("cuda:0")
    bm = torch.from_numpy(words.copy()).to(g)
    table = torch.tensor([bm.data_ptr()], dtype=torch.int64, device=g)
    slots = np.where(np.arange(n) % 3 == 0, 0, -1).astype(np.int32)
    slots[4] = 1
    ks = np.where(np.arange(n) % 2 == 0, 10, 100).astype(np.int32)
    q = (torch.from_numpy(ip).to(g), torch.from_numpy(ix.view(np.int16)).to(g), torch.from_numpy(v).to(g))
    gd, gl = dev.search_device(*q, 100, ks=torch.from_numpy(ks).to(g), params=params(nsparse, 10, 0.8),
                               selector_of_query=torch.from_numpy(slots).to(g), bitmaps=table)
    gd, gl = gd.cpu().numpy(), gl.cpu().numpy()
    assert dev.device_error() == 6
    assert (gl[4] == -1).all() and (gd[4] == -1.0).all()
    plain = {kk: o.search(ip, ix, v, kk, cut=10, hf=0.8) for kk in (10, 100)}
    masked = {kk: o.search(ip, ix, v, kk, cut=10, hf=0.8, selector=("bitmap", words, n_docs)) for kk in (10, 100)}
    for r in range(n):
        if r == 4:
            continue
        wd, wl = (masked if slots[r] == 0 else plain)[ks[r]]
        kk = ks[r]
        assert (gl[r, :kk] == wl[r]).all() and (gd[r, :kk] == wd[r]).all(), r
        assert (gl[r, kk:] == -1).all(), r
    dev.search_device(*q, 10, params=params(nsparse, 10, 0.8))
    torch.cuda.synchronize()
    assert dev.device_error() == 0


def test_idmap_with_external_id_selectors(nsparse, forced):
    """IDMapIndex over a GPU-built SEISMIC-SQ index, every search spread:
    equal to the oracle's IDMapIndex on external ids."""
    from test_gpu_search import test_idmap_search_matches_oracle

    test_idmap_search_matches_oracle(nsparse, 1)


def test_concurrent_streams_share_no_scratch(nsparse, data, pair):
    """Spread searches launched back to back on two streams, without waiting
    in between, each hold scratch of their own: every result equals the
    oracle's."""
    import torch

    o, dev, _ = pair
    g = torch.device("cuda:0")
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    batches, wants = [], []
    for qs in (data[1], data[2]):
        ip, ix, v = qs
        for i in range(0, 20, 2):
            a, b = int(ip[i]), int(ip[i + 2])
            qip = (ip[i:i + 3] - ip[i]).astype(np.int32)
            batches.append((torch.from_numpy(qip).to(g), torch.from_numpy(ix[a:b].view(np.int16)).to(g),
                            torch.from_numpy(v[a:b]).to(g)))
            wants.append(o.search(qip, ix[a:b], v[a:b], 10, cut=10, hf=0.8))
    torch.cuda.synchronize()
    p = params(nsparse, 10, 0.8)
    outs = [dev.search_device(*q, 10, params=p, stream=streams[j % 2]) for j, q in enumerate(batches)]
    torch.cuda.synchronize()
    assert dev.device_error() == 0
    for (gd, gl), want in zip(outs, wants):
        assert_same((gd.cpu().numpy(), gl.cpu().numpy()), want)
