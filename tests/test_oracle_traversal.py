"""Pins the oracle's search traversal: hand-made indexes whose results were
derived by hand from the reference's traversal (tests/_hand_cases.py, where
each derivation is written out), each built so that a plausible mistake --
break instead of continue on a later list, continue instead of break on the
first list, summaries over the cut terms only, an empty list that keeps
first_list, an out-of-range term that clears it, reversed tie order, the
wrong result-tie rule, summed repeated query terms -- changes the labels.
Also: the oracle's summary scores equal the reference's own compiled
InvertedListClusters::score_summaries_transposed bit for bit."""
import numpy as np
import pytest

import _oracle as O
from _hand_cases import CASES, build_oracle, query_arrays
from _oracle import OracleIndex


@pytest.mark.parametrize("tie_mode", [0, 1])
@pytest.mark.parametrize("name", sorted(CASES))
def test_hand_derived_traversal(name, tie_mode):
    case = CASES[name]
    o = build_oracle(case, tie_mode)
    d, lab = o.search(*query_arrays(case), case["k"], cut=case["cut"], hf=case["hf"])
    want_l, want_d = case["want"][tie_mode]
    assert list(lab[0]) == want_l, name
    assert list(d[0]) == want_d, name


def test_hand_cases_survive_the_writer():
    """Each hand index written and read back (write_index layout) searches the
    same: the GPU tests hand these bytes to the device."""
    for name, case in CASES.items():
        o = build_oracle(case, 1)
        back = OracleIndex.load(o.serialize(), tie_mode=1)
        q = query_arrays(case)
        a = o.search(*q, case["k"], cut=case["cut"], hf=case["hf"])
        b = back.search(*q, case["k"], cut=case["cut"], hf=case["hf"])
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), name


def test_first_list_break_stops_the_prune_tests():
    """The traversal counters of the first_list_break case: the break leaves
    the third cluster untested (2 prune tests, 1 cluster opened)."""
    case = CASES["first_list_break"]
    o = build_oracle(case, 0)
    _, _, st = o.search(*query_arrays(case), 1, cut=1, hf=-1.0, with_stats=True)
    stats = dict(zip(O.STAT_FIELDS, st[0]))
    assert stats["clusters_tested"] == 2 and stats["clusters_opened"] == 1 and stats["docs_scored"] == 1


def corpus(n, dim, nnz, seed):
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, dim + 1) ** 1.1
    p /= p.sum()
    ip, ix, v = [0], [], []
    for _ in range(n):
        m = int(np.clip(round(np.exp(np.log(nnz) + 0.35 * rng.standard_normal())), 2, dim))
        ix.append(np.sort(rng.choice(dim, size=m, replace=False, p=p)))
        v.append(rng.lognormal(0, 0.7, m).astype(np.float32))
        ip.append(ip[-1] + m)
    return np.array(ip, np.int32), np.concatenate(ix).astype(np.uint16), np.concatenate(v).astype(np.float32)


@pytest.mark.skipif(O.ref is None, reason="reference build-side library not built here")
@pytest.mark.parametrize("sq", [0, 1])
@pytest.mark.parametrize("n,dim,lam,beta", [(3000, 2048, 200, 20), (1500, 30522, 100, 10)])
def test_summary_scores_equal_compiled_reference(sq, n, dim, lam, beta):
    """Every query against the lists of its terms (all query entries, CSR
    order, a repeated term in some rows): the oracle's scores equal the
    reference's bit for bit."""
    ip, ix, v = corpus(n, dim, 40, seed=n + sq)
    vmin, vmax = float(v.min()), float(v.max())
    o = OracleIndex(dim, sq, lam, beta, 0.4, 3, vmin, vmax, tie_mode=0)
    o.add(ip, ix, v)
    o.build()
    qip, qix, qv = corpus(30, dim, 20, seed=n + 7)
    rows, lists = [], []
    for q in range(30):
        t, w = qix[qip[q]:qip[q + 1]], qv[qip[q]:qip[q + 1]]
        if q % 4 == 0:  # repeat the first term with another weight
            t, w = np.append(t, t[0]).astype(np.uint16), np.append(w, 1.7 * w[0]).astype(np.float32)
        for term in t[:8]:
            rows.append((t, w))
            lists.append(int(term))
    pip = np.concatenate([[0], np.cumsum([len(r[0]) for r in rows])]).astype(np.int32)
    pix = np.concatenate([r[0] for r in rows]).astype(np.uint16)
    pv = np.concatenate([r[1] for r in rows]).astype(np.float32)
    got = o.score_summaries(pip, pix, pv, lists)
    want = O.ref_score_summaries(dim, sq, (ip, ix, v), lam, beta, 0.4, 3, (pip, pix, pv), lists, vmin, vmax)
    assert sum(int((w != 0).sum()) for w in want) > 100
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))
