"""Serving: continuous batching over a device-resident index. Every request's
result must equal the oracle's search of that query on the same index (same k
and selector; scores bit for bit, labels in the id tie mode) and the device's
search of that query alone (nsparse_search_device with n = 1), whatever batch
and stream it landed in; a tight scratch budget must hold requests back
without losing any."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

DIM = 30522


@pytest.fixture(scope="module")
def served(nsparse):
    from nsparse import synth

    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=12), 20000)
    vmax = float(v.max())
    idx = nsparse.index_factory(DIM, f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda=300|beta=30|"
                                     "alpha=0.4|seed=4")
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    return idx, len(ip) - 1, O.OracleIndex.load_from_index(nsparse, idx, tie_mode=1)


def bitmap_words(n_docs, seed):
    keep = np.random.default_rng(seed).random(n_docs) >= 0.2  # excludes a random 20% of ids
    b = np.packbits(keep.astype(np.uint8), bitorder="little")
    return np.pad(b, (0, (-len(b)) % 4)).view(np.uint32), keep


def alone(nsparse, idx, qi, qv, k, words):
    import torch

    dev = torch.device("cuda:0")
    a = torch.tensor([0, len(qi)], dtype=torch.int32, device=dev)
    b = torch.from_numpy(qi.astype(np.uint16).view(np.int16)).to(dev)
    c = torch.from_numpy(qv.astype(np.float32)).to(dev)
    sel = table = bm = None
    if words is not None:
        bm = torch.from_numpy(words.copy()).to(dev)
        table = torch.tensor([bm.data_ptr()], dtype=torch.int64, device=dev)
        sel = torch.zeros(1, dtype=torch.int32, device=dev)
    d, lab = idx.search_device(a, b, c, k, params=nsparse.SeismicSearchParameters(10, 0.8),
                               selector_of_query=sel, bitmaps=table)
    torch.cuda.synchronize()
    return d.cpu().numpy()[0], lab.cpu().numpy()[0]


def test_each_result_is_the_same_alone_and_in_the_mix(nsparse, served):
    from nsparse import synth

    idx, n_docs, o = served
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=13, log_uniform=True), 400)
    n = len(qip) - 1
    rng = np.random.default_rng(14)
    ks = np.where(rng.random(n) < 0.5, 10, 100).astype(np.int32)
    pool = [bitmap_words(n_docs, 20 + i) for i in range(3)]
    bids = np.where(rng.random(n) < 0.1, rng.integers(0, 3, n), -1).astype(np.int32)
    arrivals = np.cumsum(rng.exponential(1.0 / 20000.0, n))  # a burst: 20k requests/s
    per_req = 512 * 6 + 12 + 8 * 100
    srv = nsparse.Server(idx, max_batch=24, streams=4, scratch_budget=3 * 24 * per_req, k_max=100, cut=10,
                         heap_factor=0.8)
    for words, _ in pool:
        srv.add_bitmap(words)
    d, lab, lat = srv.replay(arrivals, qip, qix, qv, ks, bids)
    st = srv.stats()
    srv.close()
    assert st["completed"] == n and st["batches"] > 1
    assert st["max_inflight_bytes"] <= 3 * 24 * per_req
    assert (lat > 0).all()
    od, ol = O.search_mixed(o, (qip, qix, qv), ks, 100, 10, 0.8, bids, [w for w, _ in pool], n_docs)
    assert np.array_equal(lab, ol) and np.array_equal(d, od)
    for i in range(n):
        a, b = qip[i], qip[i + 1]
        k = int(ks[i])
        words = pool[bids[i]][0] if bids[i] >= 0 else None
        wd, wl = alone(nsparse, idx, qix[a:b], qv[a:b], k, words)
        assert (lab[i, :k] == wl).all() and (d[i, :k] == wd).all(), i
        assert (lab[i, k:] == -1).all()
        if bids[i] >= 0:
            got = lab[i, :k][lab[i, :k] >= 0]
            assert pool[bids[i]][1][got].all()  # excluded ids never come back


def test_submit_wait_and_budget_backpressure(nsparse, served):
    from nsparse import synth

    idx, _, o = served
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=15, log_uniform=True), 120)
    od, ol = o.search(qip, qix, qv, 10, cut=10, hf=0.8)
    per_req = 512 * 6 + 12 + 8 * 10
    # room for exactly two full-length requests in flight
    srv = nsparse.Server(idx, max_batch=64, streams=4, scratch_budget=2 * per_req, k_max=10, cut=10,
                         heap_factor=0.8)
    tickets = [srv.submit(qix[qip[i]:qip[i + 1]], qv[qip[i]:qip[i + 1]], 10) for i in range(len(qip) - 1)]
    for i, t in enumerate(tickets):
        d, lab, us = srv.wait(t, 10)
        wd, wl = alone(nsparse, idx, qix[qip[i]:qip[i + 1]], qv[qip[i]:qip[i + 1]], 10, None)
        assert (lab == wl).all() and (d == wd).all() and us > 0
        assert (lab == ol[i]).all() and (d == od[i]).all()
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(tickets)
    assert st["max_inflight_bytes"] <= 2 * per_req
    assert st["deferred_by_budget"] > 0


def test_bad_requests_are_rejected(nsparse, served):
    idx, _, _ = served
    srv = nsparse.Server(idx, max_batch=8, streams=1, k_max=10)
    with pytest.raises(ValueError):
        srv.submit(np.array([1], np.uint16), np.array([1.0], np.float32), 11)  # k above k_max
    with pytest.raises(ValueError):
        srv.submit(np.array([40000 % 65536], np.uint16), np.array([1.0], np.float32), 5)  # term >= dim
    with pytest.raises(ValueError):
        srv.submit(np.zeros(600, np.uint16), np.ones(600, np.float32), 5)  # too long
    with pytest.raises(ValueError):
        srv.submit(np.array([1], np.uint16), np.array([1.0], np.float32), 5, bitmap_id=3)
    srv.close()
