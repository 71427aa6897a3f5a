"""Concurrent drop-in searches on one index, as the reference allows
(its search keeps no shared mutable state; threads share a built index).
The index is freshly read from a file, so the first searches also race to
upload it to HBM; every thread's results must equal the sequential ones."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIM = 30522


def _concurrent(index, queries, k, params, n_threads=8):
    qip, qix, qv = queries
    nq = len(qip) - 1
    out = [None] * n_threads
    errors = []
    start = threading.Barrier(n_threads)

    def run(i):
        try:
            start.wait()
            out[i] = index.search(nq, qip, qix, qv, k, params)
        except Exception as e:  # surfaced below
            errors.append(e)

    ts = [threading.Thread(target=run, args=(i,)) for i in range(n_threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    return out


@pytest.mark.parametrize("spec", ["seismic_sq", "seismic", "inverted", "idmap,inverted"])
def test_concurrent_searches_equal_sequential(nsparse, spec):
    from nsparse import synth

    print(f"concurrent searches on {spec}", flush=True)
    docs = synth.make_corpus(synth.doc_spec(seed=41), 20000)
    queries = synth.make_corpus(synth.query_spec(seed=42), 64)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    desc = {"seismic_sq": f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=200|beta=20|alpha=0.4|seed=0",
            "seismic": "seismic,lambda=200|beta=20|alpha=0.4|seed=0"}.get(spec, spec)
    idx = nsparse.index_factory(DIM, desc)
    if spec.startswith("idmap"):
        idx.add_with_ids(len(docs[0]) - 1, *docs, np.arange(20000, dtype=np.int32) * 3 + 5)
    else:
        idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    params = nsparse.SeismicSearchParameters(10, 0.8) if spec.startswith("seismic") else None
    want_d, want_l = idx.search(len(queries[0]) - 1, *queries, 10, params)
    fresh = nsparse.read_index_bytes(nsparse.serialize_index(idx))
    for d, lab in _concurrent(fresh, queries, 10, params):
        assert np.array_equal(lab, want_l) and np.array_equal(d, want_d)


@pytest.mark.parametrize("spec", ["seismic_sq", "inverted", "brutal"])
def test_reused_staging_equals_fresh_index(nsparse, spec):
    """Searches over host arrays reuse one index's staging (a stream, a device
    arena, a pinned mirror): calls that grow, shrink and switch the selector on
    one index must equal the same calls on indexes searched for the first time."""
    from nsparse import synth

    print(f"staging reuse on {spec}", flush=True)
    docs = synth.make_corpus(synth.doc_spec(seed=51), 12000)
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=52), 48)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    desc = {"seismic_sq": f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=200|beta=20|alpha=0.4|seed=0"
            }.get(spec, spec)

    def fresh():
        idx = nsparse.index_factory(DIM, desc)
        idx.add(len(docs[0]) - 1, *docs)
        if spec != "brutal":  # BrutalIndex has no build step (Index::build throws)
            idx.build()
        return idx

    def params(sel=None):
        p = nsparse.SeismicSearchParameters(10, 0.8) if spec == "seismic_sq" else (
            nsparse.SearchParameters() if sel is not None else None)
        if sel is not None:
            p.set_id_selector(sel)
        return p

    sub = lambda a, b: (qip[a:b + 1] - qip[a], qix[qip[a]:qip[b]], qv[qip[a]:qip[b]])  # noqa: E731
    few = np.array([3, 17, 4000, 9001, 11999], dtype=np.int32)
    calls = [(sub(0, 48), 50, None), (sub(5, 6), 10, None), (sub(10, 14), 10, "few"), (sub(0, 48), 20, "many"),
             (sub(7, 8), 1, None)]
    shared = fresh()
    for q, k, sel_kind in calls:
        if spec == "brutal" and sel_kind is not None:
            continue
        sel = {"few": lambda: nsparse.SetIDSelector(few),
               "many": lambda: nsparse.SetIDSelector(np.arange(0, 12000, 3, dtype=np.int32))}.get(sel_kind)
        got = shared.search(len(q[0]) - 1, *q, k, params(sel() if sel else None))
        want = fresh().search(len(q[0]) - 1, *q, k, params(sel() if sel else None))
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[0], want[0]), (spec, k, sel_kind)


@pytest.mark.parametrize("spec", ["seismic_sq", "inverted"])
def test_concurrent_mixed_sizes_equal_sequential(nsparse, spec):
    """Threads searching one index at once with batches of different sizes and
    k: each takes a stage from the index's pool (growing it as needed), and
    every result equals the same call made alone beforehand."""
    from nsparse import synth

    print(f"concurrent mixed-size searches on {spec}", flush=True)
    docs = synth.make_corpus(synth.doc_spec(seed=61), 15000)
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=62), 40)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    desc = {"seismic_sq": f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=200|beta=20|alpha=0.4|seed=0"
            }.get(spec, spec)
    idx = nsparse.index_factory(DIM, desc)
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    params = nsparse.SeismicSearchParameters(10, 0.8) if spec == "seismic_sq" else None
    sub = lambda a, b: (qip[a:b + 1] - qip[a], qix[qip[a]:qip[b]], qv[qip[a]:qip[b]])  # noqa: E731
    calls = [(sub(0, 40), 30), (sub(3, 4), 10), (sub(8, 20), 100), (sub(21, 22), 1), (sub(0, 10), 10),
             (sub(30, 40), 50)]
    want = [idx.search(len(q[0]) - 1, *q, k, params) for q, k in calls]
    n_threads = 6
    got = [[None] * len(calls) for _ in range(n_threads)]
    errors = []
    start = threading.Barrier(n_threads)

    def run(t):
        try:
            start.wait()
            for i in range(len(calls)):
                c = (i + t) % len(calls)  # each thread walks the calls from its own offset
                q, k = calls[c]
                got[t][c] = idx.search(len(q[0]) - 1, *q, k, params)
        except Exception as e:  # surfaced below
            errors.append(e)

    ts = [threading.Thread(target=run, args=(t,)) for t in range(n_threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for t in range(n_threads):
        for c, (d, lab) in enumerate(got[t]):
            assert np.array_equal(lab, want[c][1]) and np.array_equal(d, want[c][0]), (t, c)


@pytest.mark.parametrize("spec", ["seismic_sq", "inverted"])
def test_first_stage_on_its_own_stream(nsparse, spec, monkeypatch):
    """With NSPARSE_STAGE_NULL_STREAM=0 an index's first host-array stage runs
    on a non-blocking stream of its own instead of the null stream: single and
    concurrent searches equal the default's."""
    from nsparse import synth

    print(f"first stage on its own stream, {spec}", flush=True)
    docs = synth.make_corpus(synth.doc_spec(seed=61), 12000)
    queries = synth.make_corpus(synth.query_spec(seed=62), 40)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    desc = {"seismic_sq": f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=200|beta=20|alpha=0.4|seed=0"
            }.get(spec, spec)
    idx = nsparse.index_factory(DIM, desc)
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    params = nsparse.SeismicSearchParameters(10, 0.8) if spec == "seismic_sq" else None
    blob = nsparse.serialize_index(idx)
    want_d, want_l = nsparse.read_index_bytes(blob).search(len(queries[0]) - 1, *queries, 10, params)
    monkeypatch.setenv("NSPARSE_STAGE_NULL_STREAM", "0")
    own = nsparse.read_index_bytes(blob)
    d, lab = own.search(len(queries[0]) - 1, *queries, 10, params)
    assert np.array_equal(lab, want_l) and np.array_equal(d, want_d)
    for d, lab in _concurrent(own, queries, 10, params, n_threads=4):
        assert np.array_equal(lab, want_l) and np.array_equal(d, want_d)
