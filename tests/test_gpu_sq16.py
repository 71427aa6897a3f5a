"""16-bit scalar quantization (QuantizerType::QT_16bit) on the device against
the oracle: the device build of a 16-bit SEISMIC-SQ index writes the oracle's
bytes (id tie mode), and searches -- host arrays, the batched device entry with
mixed k and per-query selectors, a captured plan, the exact-match path,
selectors and a query range -- return the oracle's scores bit for bit and its
labels, on the device's own build and on a reference-order (libstdc++ tie mode)
file read from bytes. The oracle's 16-bit build is pinned against the
reference's compiled build (tests/test_oracle.py, tests/test_golden.py)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

DIM = 30522


@pytest.fixture(scope="module")
def data(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=0), 12000)
    queries = synth.make_corpus(synth.query_spec(seed=1), 120)
    mixed = synth.make_corpus(synth.query_spec(seed=2, log_uniform=True), 120)
    return docs, queries, mixed


@pytest.fixture(scope="module")
def built(nsparse, data):
    ip, ix, v = data[0]
    vmax = np.float32(v.max())
    desc = f"seismic_sq,quantizer=16bit|vmin=0|vmax={vmax:.9g}|lambda=300|beta=30|alpha=0.4|seed=0"
    print("16-bit device build", desc, flush=True)
    idx = nsparse.index_factory(DIM, desc)
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    o = O.OracleIndex(DIM, 2, 300, 30, 0.4, 0, 0.0, vmax, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    return idx, o, vmax


def assert_same(got, want):
    gd, gl = got
    wd, wl = want
    assert (gl == wl).all(), np.argwhere(gl != wl)[:5]
    assert (gd == wd).all(), np.argwhere(gd != wd)[:5]


def test_device_build_writes_the_oracle_bytes(nsparse, built):
    idx, o, _ = built
    blob = nsparse.serialize_index(idx)
    assert blob[8] == 1  # QuantizerType::QT_16bit after the fourcc and dim
    assert blob == o.serialize()


@pytest.mark.parametrize("cut,hf,k", [(10, 0.8, 10), (10, 1.0, 100), (3, 0.7, 10), (1000, 0.9, 10), (1, 0.5, 5)])
def test_search_matches_oracle(nsparse, data, built, cut, hf, k):
    idx, o, _ = built
    for ip, ix, v in data[1:]:
        print("16-bit search", cut, hf, k, flush=True)
        got = idx.search(len(ip) - 1, ip, ix, v, k, nsparse.SeismicSearchParameters(cut, hf))
        assert_same(got, o.search(ip, ix, v, k, cut=cut, hf=hf))


def test_selectors_and_query_range(nsparse, data, built):
    idx, o, vmax = built
    ip, ix, v = data[1]
    n = len(ip) - 1
    n_docs = len(data[0][0]) - 1
    rng = np.random.default_rng(7)
    allowed = np.sort(rng.choice(n_docs, n_docs // 2, replace=False)).astype(np.int32)
    banned = np.sort(rng.choice(n_docs, n_docs // 5, replace=False)).astype(np.int32)

    def p(sel=None):
        q = nsparse.SeismicSearchParameters(10, 0.9)
        if sel is not None:
            q.set_id_selector(sel)
        return q

    print("16-bit selectors", flush=True)
    assert_same(idx.search(n, ip, ix, v, 10, p(nsparse.SetIDSelector(allowed))),
                o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("set", allowed)))
    assert_same(idx.search(n, ip, ix, v, 10, p(nsparse.NotIDSelector(nsparse.SetIDSelector(banned)))),
                o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("set", banned), negate=True))
    mask = np.zeros(n_docs, bool)
    mask[allowed] = True
    bm = nsparse.BitmapIDSelector(mask)
    assert_same(idx.search(n, ip, ix, v, 10, p(bm)),
                o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("bitmap", bm.words, n_docs)))
    # size <= k: the exact-match path, a repeated id included
    few = np.array([allowed[3], allowed[1], allowed[3], allowed[7]], np.int32)
    assert_same(idx.search(n, ip, ix, v, 10, p(nsparse.ArrayIDSelector(few))),
                o.search(ip, ix, v, 10, cut=10, hf=0.9, selector=("array", few)))
    # SeismicSQSearchParameters: the query's own range, the index's width
    rq = nsparse.SeismicSQSearchParameters(0.0, float(2 * vmax), 10, 0.8)
    assert_same(idx.search(n, ip, ix, v, 10, rq),
                o.search(ip, ix, v, 10, cut=10, hf=0.8, sq_range=(0.0, np.float32(2 * vmax))))


def test_batched_device_and_plan(nsparse, data, built):
    import torch

    idx, o, _ = built
    ip, ix, v = data[2]
    n = len(ip) - 1
    n_docs = len(data[0][0]) - 1
    ks = np.where(np.arange(n) % 2 == 0, 10, 100).astype(np.int32)
    dev = torch.device("cuda:0")
    d_ip = torch.from_numpy(ip).to(dev)
    d_ix = torch.from_numpy(ix.view(np.int16)).to(dev)
    d_v = torch.from_numpy(v).to(dev)
    d_ks = torch.from_numpy(ks).to(dev)
    rng = np.random.default_rng(9)
    mask = rng.random(n_docs) >= 0.2
    words = np.packbits(mask.astype(np.uint8), bitorder="little").view(np.uint8)
    words = np.pad(words, (0, (-len(words)) % 4)).view(np.uint32)
    bm = torch.from_numpy(words.copy()).to(dev)
    table = torch.tensor([bm.data_ptr()], dtype=torch.int64, device=dev)
    sel = np.where(np.arange(n) % 7 == 0, 0, -1).astype(np.int32)
    d_sel = torch.from_numpy(sel).to(dev)
    params = nsparse.SeismicSearchParameters(10, 0.8)
    want = O.search_mixed(o, (ip, ix, v), ks, 100, 10, 0.8, sel, [words], n_docs)
    print("16-bit batched device search", flush=True)
    dd, dl = idx.search_device(d_ip, d_ix, d_v, 100, ks=d_ks, params=params, selector_of_query=d_sel, bitmaps=table)
    torch.cuda.synchronize()
    assert idx.device_error() == 0
    assert_same((dd.cpu().numpy(), dl.cpu().numpy()), want)
    # traversal counters: the oracle's, rows re-scored for docs seen in an
    # earlier list aside (the device checks those against its top-k)
    stats = torch.zeros((n, 16), dtype=torch.int64, device=dev)
    idx.search_device(d_ip, d_ix, d_v, 100, ks=d_ks, params=params, stats=stats)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    w10 = o.search(ip, ix, v, 10, cut=10, hf=0.8, with_stats=True)[2]
    w100 = o.search(ip, ix, v, 100, cut=10, hf=0.8, with_stats=True)[2]
    for q in range(n):
        ws = w10 if ks[q] == 10 else w100
        assert (st[q, :6] == ws[q, :6]).all(), (q, st[q], ws[q])
        assert st[q, 6] >= ws[q, 6]
    for cfg in (dict(lookup=0, block=256, persistent=1, longest_first=1), None):
        print("16-bit plan", cfg, flush=True)
        plan = idx.search_plan(d_ip, d_ix, d_v, 100, ks=d_ks, params=params, config=cfg, selector_of_query=d_sel,
                               bitmaps=table)
        for _ in range(2):
            plan.out_l.fill_(7)
            torch.cuda.synchronize()
            got = plan.launch()
            torch.cuda.synchronize()
            assert_same(tuple(t.cpu().numpy() for t in got), want)
        plan.free()
    assert idx.device_error() == 0


def test_reference_order_file(nsparse, data):
    """A 16-bit file in the reference's own layout (the oracle's libstdc++ tie
    mode, byte-identical to the reference's build) read from bytes: the device
    search equals the oracle's id-tie search of the same index."""
    ip, ix, v = data[0]
    vmax = np.float32(v.max())
    o0 = O.OracleIndex(DIM, 2, 300, 30, 0.4, 5, 0.0, vmax, tie_mode=0)
    o0.add(ip, ix, v)
    o0.build()
    blob = o0.serialize()
    dev = nsparse.read_index_bytes(blob)
    assert nsparse.serialize_index(dev) == blob
    o1 = O.OracleIndex.load(blob, tie_mode=1)
    qip, qix, qv = data[2]
    print("16-bit reference-order search", flush=True)
    got = dev.search(len(qip) - 1, qip, qix, qv, 10, nsparse.SeismicSearchParameters(10, 0.8))
    assert_same(got, o1.search(qip, qix, qv, 10, cut=10, hf=0.8))


def test_serving_loop(nsparse, data, built):
    """The continuous-batching server over the 16-bit index: a Poisson burst of
    mixed-k requests, 10% with a 20%-exclusion bitmap, equals the oracle."""
    idx, o, _ = built
    ip, ix, v = data[2]
    n = len(ip) - 1
    n_docs = len(data[0][0]) - 1
    rng = np.random.default_rng(14)
    ks = np.where(rng.random(n) < 0.5, 10, 100).astype(np.int32)
    keep = rng.random(n_docs) >= 0.2
    words = np.packbits(keep.astype(np.uint8), bitorder="little")
    words = np.pad(words, (0, (-len(words)) % 4)).view(np.uint32)
    bids = np.where(rng.random(n) < 0.1, 0, -1).astype(np.int32)
    arrivals = np.cumsum(rng.exponential(1.0 / 20000.0, n))
    print("16-bit serving replay", flush=True)
    srv = nsparse.Server(idx, max_batch=24, streams=4, k_max=100, cut=10, heap_factor=0.8)
    srv.add_bitmap(words)
    d, lab, _ = srv.replay(arrivals, ip, ix, v, ks, bids)
    srv.close()
    want = O.search_mixed(o, (ip, ix, v), ks, 100, 10, 0.8, bids, [words], n_docs)
    assert_same((d, lab), want)
