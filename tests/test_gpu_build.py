"""GPU build parity: add() + build() on the device, then write_index, must give
the oracle's bytes exactly (id tie mode: value ties broken by doc id / term
id). For FP32 data, where values do not tie, that is also the reference's
own layout: the device's bytes hash to the digests of the reference's own
build in tests/golden/ref_fixtures.json."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import _oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from golden_cases import BUILD_CASES, build_corpus  # noqa: E402

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_fixtures.json")) as _f:
    FIX = json.load(_f)

pytestmark = pytest.mark.gpu

DIM = 30522


def gpu_build(nsparse, desc, docs):
    ip, ix, v = docs
    idx = nsparse.index_factory(DIM, desc)
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    return idx


def oracle_bytes(docs, sq, lam, beta, alpha, seed, vmin=0.0, vmax=1.0, tie_mode=1):
    ip, ix, v = docs
    o = O.OracleIndex(DIM, sq, lam, beta, alpha, seed, vmin, vmax, tie_mode=tie_mode)
    o.add(ip, ix, v)
    o.build()
    return o.serialize()


@pytest.mark.parametrize("n,lam,beta", [(3000, 200, 20), (12000, 400, 40), (2000, -1, -1), (1500, 50, 0)])
def test_sq_build_matches_oracle(nsparse, n, lam, beta):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=n), n)
    vmax = float(docs[2].max())
    desc = f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda={lam}|beta={beta}|alpha=0.4|seed=5"
    got = nsparse.serialize_index(gpu_build(nsparse, desc, docs))
    want = oracle_bytes(docs, 1, lam, beta, 0.4, 5, 0.0, np.float32(vmax))
    assert len(got) == len(want)
    assert got == want


@pytest.mark.parametrize("n,lam,beta,alpha", [(3000, 200, 20, 0.4), (6000, 500, 50, 0.5), (1000, 100, 0, 0.3)])
def test_f32_build_matches_oracle(nsparse, n, lam, beta, alpha):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=100 + n), n)
    desc = f"seismic,lambda={lam}|beta={beta}|alpha={alpha}|seed=9"
    got = nsparse.serialize_index(gpu_build(nsparse, desc, docs))
    want = oracle_bytes(docs, 0, lam, beta, alpha, 9)
    assert got == want


F32_GOLDEN = [i for i, c in enumerate(BUILD_CASES) if c["sq"] == 0]


@pytest.mark.parametrize("i", F32_GOLDEN, ids=[f"n{BUILD_CASES[i]['n']}-l{BUILD_CASES[i]['lam']}" for i in F32_GOLDEN])
def test_f32_build_hashes_to_reference(nsparse, i):
    """FP32 corpora whose values never tie in a way that reaches the layout:
    the device's index bytes are the reference's own (SHA-256 and size of the
    reference build's write_index output, kept in the golden fixtures)."""
    case, fx = BUILD_CASES[i], FIX["build"][i]
    ip, ix, v = build_corpus(case)
    idx = nsparse.index_factory(case["dim"], f"seismic,lambda={case['lam']}|beta={case['beta']}|"
                                             f"alpha={case['alpha']}|seed={case['seed']}")
    idx.add(len(ip) - 1, ip, ix, v)
    idx.build()
    got = nsparse.serialize_index(idx)
    assert (hashlib.sha256(got).hexdigest(), len(got)) == (fx["sha256"], fx["size"])


def test_build_then_search_matches_oracle(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=21), 8000)
    vmax = float(docs[2].max())
    desc = f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda=300|beta=30|alpha=0.4|seed=1"
    dev = gpu_build(nsparse, desc, docs)
    o = O.OracleIndex(DIM, 1, 300, 30, 0.4, 1, 0.0, np.float32(vmax), tie_mode=1)
    o.add(*docs)
    o.build()
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=22), 100)
    want = o.search(qip, qix, qv, 10, cut=10, hf=0.8)
    got = dev.search(len(qip) - 1, qip, qix, qv, 10, nsparse.SeismicSearchParameters(10, 0.8))
    assert (got[1] == want[1]).all() and (got[0] == want[0]).all()


def test_same_seed_same_index_different_seed_differs(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=31), 2000)
    a = nsparse.serialize_index(gpu_build(nsparse, "seismic,lambda=100|beta=10|alpha=0.4|seed=3", docs))
    b = nsparse.serialize_index(gpu_build(nsparse, "seismic,lambda=100|beta=10|alpha=0.4|seed=3", docs))
    c = nsparse.serialize_index(gpu_build(nsparse, "seismic,lambda=100|beta=10|alpha=0.4|seed=4", docs))
    assert a == b and a != c


@pytest.mark.parametrize("sq", [1, 0])
def test_build_in_many_list_batches_matches_oracle(nsparse, monkeypatch, sq):
    """The centroid index and the summary transpose are built in batches of
    consecutive posting lists (32-bit sort values per batch). A tiny batch cap
    forces hundreds of batches, including lists larger than the cap; the bytes
    must not change."""
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=41), 4000)
    vmax = float(docs[2].max())
    if sq:
        desc = f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda=250|beta=25|alpha=0.4|seed=2"
    else:
        desc = "seismic,lambda=250|beta=25|alpha=0.4|seed=2"
    monkeypatch.setenv("NSPARSE_BUILD_BATCH_ENTRIES", "5000")
    got = nsparse.serialize_index(gpu_build(nsparse, desc, docs))
    monkeypatch.delenv("NSPARSE_BUILD_BATCH_ENTRIES")
    want = oracle_bytes(docs, sq, 250, 25, 0.4, 2, 0.0, np.float32(vmax) if sq else 1.0)
    assert got == want


@pytest.mark.parametrize("sq", [1, 0])
@pytest.mark.parametrize("lam,beta", [(100000, 7), (150, 20)])
def test_build_counters(nsparse, sq, lam, beta):
    """The build's counters against closed forms: kept postings are
    sum_t min(count_t, lambda); clusters are sum_t min(beta, kept_t); without
    pruning every doc sits in nnz_d lists, so the kept rows hold sum_d nnz_d^2
    non-zeros; summary entries and assignment pairs are counted."""
    from nsparse import synth

    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=13), 3000)
    vmax = float(v.max())
    spec = (f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda={lam}|beta={beta}|alpha=0.4|seed=1" if sq
            else f"seismic,lambda={lam}|beta={beta}|alpha=0.4|seed=1")
    index = nsparse.index_factory(DIM, spec)
    index.add(len(ip) - 1, ip, ix, v)
    index.build()
    st = index.build_stats()
    cnt = np.bincount(ix.astype(np.int64), minlength=DIM)
    kept = np.minimum(cnt, lam)
    assert st["kept_postings"] == int(kept.sum())
    assert st["clusters"] == int(np.minimum(kept, beta).sum())
    if lam >= cnt.max():
        assert st["kept_row_nnz"] == int((np.diff(ip).astype(np.int64) ** 2).sum())
    else:
        assert 0 < st["kept_row_nnz"] < int((np.diff(ip).astype(np.int64) ** 2).sum())
    assert st["summary_entries"] > 0 and st["assign_pairs"] > 0
