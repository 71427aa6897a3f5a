"""Batched search plans (hipGraph replay + autotuned schedule): every schedule
must give exactly the oracle's results on the same index (scores bit for bit,
labels in the id tie mode), on every replay, for mixed k and per-query bitmap
selectors -- as must nsparse_search_device; query buffers rewritten in place
between replays must be searched afresh."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

DIM = 30522
CONFIGS = [dict(lookup=lk, block=b, persistent=p, longest_first=lf)
           for lk in (1, 0) for b in ((1024, 512, 256) if lk else (512, 256))
           for p, lf in ((0, 0), (1, 0), (1, 1))]


@pytest.fixture(scope="module")
def setup(nsparse):
    import torch

    from nsparse import synth

    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=5), 20000)
    vmax = float(v.max())
    idx = nsparse.index_factory(DIM, f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda=300|beta=30|"
                                     "alpha=0.4|seed=3")
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=6, log_uniform=True), 300)
    n = len(qip) - 1
    dev = torch.device("cuda:0")
    q = dict(ip=torch.from_numpy(qip).to(dev), ix=torch.from_numpy(qix.view(np.int16)).to(dev),
             v=torch.from_numpy(qv).to(dev))
    ks = torch.from_numpy(np.where(np.arange(n) % 3 == 0, 100, 10).astype(np.int32)).to(dev)
    # per-query selectors: every 7th query excludes a random 20% of the docs
    rng = np.random.default_rng(9)
    mask = rng.random(len(ip) - 1) >= 0.2
    words = np.packbits(mask.astype(np.uint8), bitorder="little").view(np.uint8)
    words = np.pad(words, (0, (-len(words)) % 4)).view(np.uint32)
    bm = torch.from_numpy(words.copy()).to(dev)
    table = torch.tensor([bm.data_ptr()], dtype=torch.int64, device=dev)
    sel = torch.from_numpy(np.where(np.arange(n) % 7 == 0, 0, -1).astype(np.int32)).to(dev)
    params = nsparse.SeismicSearchParameters(10, 0.8)
    o = O.OracleIndex.load_from_index(nsparse, idx, tie_mode=1)
    want = O.search_mixed(o, (qip, qix, qv), ks.cpu().numpy(), 100, 10, 0.8, sel.cpu().numpy(), [words],
                          len(ip) - 1)
    dev_d, dev_l = idx.search_device(q["ip"], q["ix"], q["v"], 100, ks=ks, params=params, selector_of_query=sel,
                                     bitmaps=table)
    torch.cuda.synchronize()
    assert np.array_equal(dev_l.cpu().numpy(), want[1]) and np.array_equal(dev_d.cpu().numpy(), want[0])
    keep = (bm, o, (qip, qix, qv))
    return idx, q, ks, sel, table, params, list(want), keep


def same(got, want):
    gd, gl = (t.cpu().numpy() for t in got)
    assert (gl == want[1]).all(), np.argwhere(gl != want[1])[:5]
    assert (gd == want[0]).all()


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "-".join(str(v) for v in c.values()))
def test_every_schedule_equals_device_search(setup, cfg):
    import torch

    idx, q, ks, sel, table, params, want, _ = setup
    plan = idx.search_plan(q["ip"], q["ix"], q["v"], 100, ks=ks, params=params, config=cfg, selector_of_query=sel,
                           bitmaps=table)
    for _ in range(3):  # replays, with eager work and syncs in between
        plan.out_d.fill_(7.0)
        plan.out_l.fill_(7)
        torch.cuda.synchronize()
        same(plan.launch(), want)
        torch.cuda.synchronize()
        (torch.ones(1 << 20, device="cuda") * 2).sum().item()
    assert idx.device_error() == 0
    plan.free()


def test_autotuned_plan(setup):
    import torch

    idx, q, ks, sel, table, params, want, _ = setup
    plan = idx.search_plan(q["ip"], q["ix"], q["v"], 100, ks=ks, params=params, selector_of_query=sel,
                           bitmaps=table)
    tried = plan.tuning
    assert len(tried) >= 6 and all(ms > 0 for _, ms in tried)
    assert plan.config in [c for c, _ in tried]
    assert min(ms for _, ms in tried) == dict((tuple(c.values()), ms) for c, ms in tried)[tuple(plan.config.values())]
    plan.out_l.fill_(7)
    same(plan.launch(), want)
    torch.cuda.synchronize()
    plan.free()


def test_rewritten_queries_are_searched_afresh(setup):
    import torch

    idx, q, ks, sel, table, params, want, keep = setup
    o, (qip, qix, qv) = keep[1], keep[2]
    v2 = q["v"].clone()
    plan = idx.search_plan(q["ip"], q["ix"], v2, 100, ks=ks, params=params,
                           config=dict(lookup=1, block=512, persistent=1, longest_first=1))
    base = O.search_mixed(o, (qip, qix, qv), ks.cpu().numpy(), 100, 10, 0.8)
    same(plan.launch(), base)
    # new weights in the same buffer: reversed order within the batch
    v2.copy_(torch.flip(q["v"], dims=[0]))
    torch.cuda.synchronize()
    fresh = O.search_mixed(o, (qip, qix, qv[::-1].copy()), ks.cpu().numpy(), 100, 10, 0.8)
    same(plan.launch(), fresh)
    plan.free()


def test_fp32_plan(nsparse):
    import torch

    from nsparse import synth

    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=8), 5000)
    idx = nsparse.index_factory(DIM, "seismic,lambda=200|beta=20|alpha=0.4|seed=1")
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=9), 100)
    dev = torch.device("cuda:0")
    a = torch.from_numpy(qip).to(dev)
    b = torch.from_numpy(qix.view(np.int16)).to(dev)
    c = torch.from_numpy(qv).to(dev)
    p = nsparse.SeismicSearchParameters(10, 0.9)
    want = O.OracleIndex.load_from_index(nsparse, idx, tie_mode=1).search(qip, qix, qv, 10, cut=10, hf=0.9)
    same(idx.search_device(a, b, c, 10, params=p), want)
    for cfg in (None, dict(lookup=1, block=512, persistent=1, longest_first=1)):
        plan = idx.search_plan(a, b, c, 10, params=p, config=cfg)
        assert plan.config["lookup"] == 0 and plan.config["block"] == 256
        same(plan.launch(), want)
        plan.free()
