"""The allocation-free batched search (nsparse_search_device_noalloc, through
the C ABI): its scratch is a static device array of the library, so its
results must equal nsparse_search_device's bit for bit -- on two calls in a
row, the second with fewer queries (the array keeps the first call's
contents) -- with mixed k and per-query bitmap selectors, for 8-bit, 16-bit
and FP32 indexes. Sizes beyond its limit and indexes not yet resident on the device
are rejected before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIM = 30522
MAX_QUERIES = 1 << 20


def _bitmap(n_docs, seed):
    mask = np.random.default_rng(seed).random(n_docs) >= 0.2
    words = np.packbits(mask.astype(np.uint8), bitorder="little")
    return np.pad(words, (0, (-len(words)) % 4)).view(np.uint32)


@pytest.fixture(scope="module", params=[1, 0, 2], ids=["sq", "fp32", "sq16"])
def setup(nsparse, request):
    import torch

    from nsparse import synth

    sq = request.param
    print(f"no-alloc setup sq={sq}", flush=True)
    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=21), 20000)
    vmax = float(v.max())
    width = "16bit" if sq == 2 else "8bit"
    desc = (f"seismic_sq,quantizer={width}|vmin=0|vmax={vmax:.9g}|lambda=300|beta=30|alpha=0.4|seed=2" if sq
            else "seismic,lambda=300|beta=30|alpha=0.4|seed=2")
    idx = nsparse.index_factory(DIM, desc)
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=22, log_uniform=True), 400)
    n = len(qip) - 1
    g = torch.device("cuda:0")
    bm = torch.from_numpy(_bitmap(len(ip) - 1, 23).copy()).to(g)
    table = torch.tensor([bm.data_ptr()], dtype=torch.int64, device=g)
    q = dict(ip=torch.from_numpy(qip).to(g), ix=torch.from_numpy(qix.view(np.int16)).to(g),
             v=torch.from_numpy(qv).to(g),
             ks=torch.from_numpy(np.where(np.arange(n) % 3 == 0, 100, 10).astype(np.int32)).to(g),
             sel=torch.from_numpy(np.where(np.arange(n) % 7 == 0, 0, -1).astype(np.int32)).to(g))
    return idx, q, table, bm, (qip, qix, qv)


def _expected(idx, q, table, n, p):
    import torch

    d, lab = idx.search_device(q["ip"][: n + 1], q["ix"], q["v"], 100, ks=q["ks"][:n], params=p,
                               selector_of_query=q["sel"][:n], bitmaps=table)
    torch.cuda.synchronize()
    assert idx.device_error() == 0
    return d.cpu().numpy(), lab.cpu().numpy()


def test_two_calls_equal_search_device(nsparse, setup):
    import torch

    idx, q, table, _, _ = setup
    p = nsparse.SeismicSearchParameters(10, 0.8)
    g = q["ip"].device
    err = torch.full((1,), -7, dtype=torch.int32, device=g)
    n_all = q["ip"].numel() - 1
    for n in (n_all, 37):  # the second call is smaller: stale scratch past n must not matter
        print(f"no-alloc call n={n}", flush=True)
        want = _expected(idx, q, table, n, p)
        out_d = torch.full((n, 100), 7.0, dtype=torch.float32, device=g)
        out_l = torch.full((n, 100), 7, dtype=torch.int32, device=g)
        idx.search_device_noalloc(q["ip"][: n + 1], q["ix"], q["v"], 100, out_d, out_l, err, ks=q["ks"][:n],
                                  params=p, selector_of_query=q["sel"][:n], bitmaps=table)
        torch.cuda.synchronize()
        assert int(err.item()) == 0
        assert np.array_equal(out_l.cpu().numpy(), want[1])
        assert np.array_equal(out_d.cpu().numpy(), want[0])


def test_on_a_caller_stream(nsparse, setup):
    import torch

    idx, q, table, _, _ = setup
    p = nsparse.SeismicSearchParameters(3, 0.9)
    n = q["ip"].numel() - 1
    want = _expected(idx, q, table, n, p)
    s = torch.cuda.Stream()
    g = q["ip"].device
    err = torch.zeros(1, dtype=torch.int32, device=g)
    out_d = torch.empty((n, 100), dtype=torch.float32, device=g)
    out_l = torch.empty((n, 100), dtype=torch.int32, device=g)
    torch.cuda.synchronize()
    idx.search_device_noalloc(q["ip"], q["ix"], q["v"], 100, out_d, out_l, err, ks=q["ks"], params=p,
                              selector_of_query=q["sel"], bitmaps=table, stream=s)
    s.synchronize()
    assert int(err.item()) == 0
    assert np.array_equal(out_l.cpu().numpy(), want[1]) and np.array_equal(out_d.cpu().numpy(), want[0])


def test_rejections_before_launch(nsparse, setup):
    """More queries than the array holds, and an index read from bytes that was
    never searched (not resident yet), raise ValueError; nothing is launched,
    so the outputs keep their contents."""
    import torch

    from nsparse import _lib

    idx, q, table, _, _ = setup
    g = q["ip"].device
    err = torch.full((1,), -7, dtype=torch.int32, device=g)
    out_d = torch.full((4, 10), 3.0, dtype=torch.float32, device=g)
    out_l = torch.full((4, 10), 3, dtype=torch.int32, device=g)
    rc = _lib.lib.nsparse_search_device_noalloc(idx.handle, MAX_QUERIES + 1, q["ip"].data_ptr(), q["ix"].data_ptr(),
                                                q["v"].data_ptr(), None, 10, None, None, None, 0,
                                                out_d.data_ptr(), out_l.data_ptr(), err.data_ptr(),
                                                ctypes.c_void_p(0))
    assert rc == 1 and b"at most 1048576 queries" in _lib.lib.nsparse_last_error()
    fresh = nsparse.read_index_bytes(nsparse.serialize_index(idx))
    with pytest.raises(ValueError, match="resident on the device"):
        fresh.search_device_noalloc(q["ip"][:5], q["ix"], q["v"], 10, out_d, out_l, err)
    torch.cuda.synchronize()
    assert int(err.item()) == -7 and (out_l == 3).all().item() and (out_d == 3.0).all().item()
    # after one search the copy is resident and the call goes through
    qip, qix, qv = setup[4]
    fresh.search(4, qip[:5], qix, qv, 10)
    fresh.search_device_noalloc(q["ip"][:5], q["ix"], q["v"], 10, out_d, out_l, err)
    want = idx.search_device(q["ip"][:5], q["ix"], q["v"], 10)
    torch.cuda.synchronize()
    assert int(err.item()) == 0
    assert np.array_equal(out_l.cpu().numpy(), want[1].cpu().numpy())
