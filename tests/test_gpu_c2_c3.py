"""The batched and serving workloads at full size, against the oracle:
  * C2 -- SEISMIC-SQ over 8,841,823 synthetic docs (vocab 30522, Zipf 1.1,
    doc nnz lognormal median 120, uint8 over the corpus min/max), lambda 6000,
    beta 600, alpha 0.4, built on the device; 10k queries with nnz log-uniform
    over 5..300 and k in {10, 100} at random, cut 10, heap_factor 0.8, in one
    autotuned batched plan replayed from its hipGraph. Every row equals the
    oracle's search on the device-built index bytes (scores bit for bit,
    labels in the id tie mode).
  * C3 -- the serving loop on that resident index: Poisson arrivals at 10k
    requests/s, k in {10, 100}, 10% of requests carrying a bitmap that
    excludes a random 20% of ids, under a scratch budget small enough to
    defer requests. Every request's result equals the oracle's.
The workload is the benchmark's (same generator seeds)."""
import os

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

DIM = 30522
N_DOCS = 8_841_823
N_QUERIES = 10_000
THREADS = min(16, os.cpu_count() or 1)


def _log(msg):
    print(f"[c2/c3] {msg}", flush=True)


@pytest.fixture(scope="module")
def c2(nsparse):
    from nsparse import synth

    _log("generating 8.8M docs")
    docs = synth.make_corpus(synth.doc_spec(seed=0), N_DOCS, threads=THREADS)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    idx = nsparse.index_factory(DIM, f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=6000|"
                                     "beta=600|alpha=0.4|seed=0")
    _log("device add + build")
    idx.add(N_DOCS, *docs)
    idx.build()
    del docs
    _log("oracle copy of the index")
    o = O.OracleIndex.load_from_index(nsparse, idx, tie_mode=1)
    queries = synth.make_corpus(synth.query_spec(seed=1, log_uniform=True), N_QUERIES, threads=THREADS)
    ks = np.where(np.random.default_rng(2).random(N_QUERIES) < 0.5, 10, 100).astype(np.int32)
    return idx, o, queries, ks


def test_c2_autotuned_plan_equals_oracle(nsparse, c2):
    import torch

    idx, o, (qip, qix, qv), ks = c2
    g = torch.device("cuda:0")
    q = (torch.from_numpy(qip).to(g), torch.from_numpy(qix.view(np.int16)).to(g), torch.from_numpy(qv).to(g))
    d_ks = torch.from_numpy(ks).to(g)
    _log("autotuned plan over 10k queries")
    plan = idx.search_plan(*q, 100, ks=d_ks, params=nsparse.SeismicSearchParameters(10, 0.8))
    plan.launch()
    torch.cuda.synchronize()
    assert idx.device_error() == 0
    plan.out_l.fill_(7)
    got = [t.cpu().numpy() for t in plan.launch()]  # a second replay
    _log(f"schedule {plan.config}; oracle search of 10k queries on {THREADS} threads")
    want = O.search_mixed(o, (qip, qix, qv), ks, 100, 10, 0.8, threads=THREADS)
    bad = np.nonzero((got[1] != want[1]).any(axis=1) | (got[0] != want[0]).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ, first {bad[:5]}"
    plan.free()


def test_c3_serving_equals_oracle(nsparse, c2):
    from nsparse import synth

    idx, o, _, _ = c2
    n = 4000
    rng = np.random.default_rng(3)
    pool = []
    for _ in range(4):
        keep = rng.random(N_DOCS) >= 0.2
        b = np.packbits(keep.astype(np.uint8), bitorder="little")
        pool.append(np.pad(b, (0, (-len(b)) % 4)).view(np.uint32))
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=110, log_uniform=True), n, threads=THREADS)
    r2 = np.random.default_rng(10000)
    ks = np.where(r2.random(n) < 0.5, 10, 100).astype(np.int32)
    bids = np.where(r2.random(n) < 0.1, r2.integers(0, len(pool), n), -1).astype(np.int32)
    arrivals = np.cumsum(r2.exponential(1.0 / 50000.0, n))
    per_req = 512 * 6 + 12 + 8 * 100
    # room for about 10 average requests: arrivals at 50k/s, each served in
    # 0.2 ms or more, outrun it
    budget = 4 * per_req
    srv = nsparse.Server(idx, max_batch=64, streams=4, scratch_budget=budget, k_max=100, cut=10, heap_factor=0.8)
    for w in pool:
        srv.add_bitmap(w)
    _log(f"serving replay: {n} requests at 50k/s")
    d, lab, lat = srv.replay(arrivals, qip, qix, qv, ks, bids)
    st = srv.stats()
    srv.close()
    _log(f"served: {st}; p50 {np.median(lat) / 1e3:.2f} ms")
    assert st["completed"] == n and st["max_inflight_bytes"] <= budget
    assert st["deferred_by_budget"] > 0
    want = O.search_mixed(o, (qip, qix, qv), ks, 100, 10, 0.8, bids, pool, N_DOCS, threads=THREADS)
    bad = np.nonzero((lab != want[1]).any(axis=1) | (d != want[0]).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} requests differ, first {bad[:5]}"
