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
