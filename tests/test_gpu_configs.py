"""The reference's own shapes beside the batched benchmark:
  * C0 -- FP32 SEISMIC, 100k docs, lambda=2000, centroid fraction 0.1 (beta=-1),
    alpha=0.4, seed 0; 1k queries (nnz lognormal median 40), cut 10, hf 0.8,
    k 10: the GPU build must write the oracle's bytes and the GPU search must
    return the oracle's results;
  * C4 -- differentiable re-scoring: 256 FP32 queries x 1000 candidates mined
    by SEISMIC from a 1M-doc FP32 index. Forward scores against float64 sparse
    products on the CPU; the backward against a CPU reference gradient; the
    InfoNCE gradient (temperature 0.05, first candidate positive) against FP64
    central finite differences (eps 1e-3, relative error <= 1e-4)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

DIM = 30522


def test_c0_fp32_100k_build_and_search(nsparse):
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=0), 100_000)
    queries = synth.make_corpus(synth.query_spec(seed=1), 1000)
    desc = "seismic,lambda=2000|beta=-1|alpha=0.4|seed=0"
    idx = nsparse.index_factory(DIM, desc)
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    got_bytes = nsparse.serialize_index(idx)
    o = O.OracleIndex(DIM, 0, 2000, 200, 0.4, 0, tie_mode=1)  # beta = floor(0.1 * 2000)
    o.add(*docs)
    o.build()
    assert got_bytes == o.serialize()
    want = o.search(*queries, 10, cut=10, hf=0.8, threads=16)
    got = idx.search(len(queries[0]) - 1, *queries, 10, nsparse.SeismicSearchParameters(10, 0.8))
    assert (got[1] == want[1]).all() and (got[0] == want[0]).all()


def test_c1_sq_1m_plain_search_and_recall(nsparse):
    """C1 -- SEISMIC-SQ over 1M docs (per-index min/max), lambda=6000, beta=600
    (centroid fraction 0.1), alpha=0.4, cut 10: 10k queries (C0's distributions)
    at k=10 one per launch through the plain drop-in at heap_factor 0.8, and
    batched at 0.7 and 0.9. Every result equals the oracle's on the device-built
    index (scores bit for bit). Recall@10 against exact FP32 brute force over 200
    queries is then the oracle's recall exactly, and clears the reference python
    tests' SQ floor (test_seismic_sq_index.py:46, 0.82)."""
    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=0), 1_000_000, threads=16)
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=1), 10_000)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    idx = nsparse.index_factory(DIM, f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=6000|"
                                     "beta=600|alpha=0.4|seed=0")
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    o = O.OracleIndex.load_from_index(nsparse, idx, tie_mode=1)
    nq = len(qip) - 1
    print("C1: 10k plain searches at hf 0.8", flush=True)
    p = nsparse.SeismicSearchParameters(10, 0.8)
    got_d = np.empty((nq, 10), np.float32)
    got_l = np.empty((nq, 10), np.int32)
    for i in range(nq):
        a, b = int(qip[i]), int(qip[i + 1])
        got_d[i], got_l[i] = (x[0] for x in idx.search(1, np.array([0, b - a], np.int32), qix[a:b], qv[a:b], 10, p))
    want_d, want_l = o.search(qip, qix, qv, 10, cut=10, hf=0.8, threads=16)
    assert np.array_equal(got_l, want_l) and np.array_equal(got_d, want_d)
    for hf in (0.7, 0.9):
        d, lab = idx.search(nq, qip, qix, qv, 10, nsparse.SeismicSearchParameters(10, hf))
        wd, wl = o.search(qip, qix, qv, 10, cut=10, hf=hf, threads=16)
        assert np.array_equal(lab, wl) and np.array_equal(d, wd), hf
    # recall@10 of the plain results against exact FP32 scores on the unquantized docs
    D = _csr(*docs)
    hits = []
    for q0 in range(0, 200, 20):
        Q = _csr(qip[q0:q0 + 21] - qip[q0], qix[qip[q0]:qip[q0 + 20]], qv[qip[q0]:qip[q0 + 20]])
        S = (Q @ D.T).toarray()
        for r in range(S.shape[0]):
            (el,) = np.nonzero(S[r] > 0)
            truth = set(el[np.lexsort((el, -S[r][el]))][:10].tolist())
            if truth:
                hits.append(len(truth & set(got_l[q0 + r].tolist())) / len(truth))
    recall = float(np.mean(hits))
    print(f"C1 recall@10 (hf 0.8, 200 queries) = {recall:.4f}", flush=True)
    assert recall >= 0.82, recall


def _csr(ip, ix, v, n_cols=DIM):
    import scipy.sparse as sp

    m = sp.csr_matrix((v.astype(np.float64), ix.astype(np.int64), ip.astype(np.int64)),
                      shape=(len(ip) - 1, n_cols))
    assert m.has_sorted_indices and m.nnz == len(v)  # entries stay in the given order
    return m


@pytest.fixture(scope="module")
def c4(nsparse):
    import torch

    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=0), 1_000_000, threads=16)
    queries = synth.make_corpus(synth.query_spec(seed=1), 256)
    idx = nsparse.index_factory(DIM, "seismic,lambda=6000|beta=600|alpha=0.4|seed=0")
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    dev = torch.device("cuda")
    t = dict(q_indptr=torch.from_numpy(queries[0]).to(dev),
             q_indices=torch.from_numpy(queries[1].view(np.int16)).to(dev),
             d_indptr=torch.from_numpy(docs[0]).to(dev),
             d_indices=torch.from_numpy(docs[1].view(np.int16)).to(dev))
    _, cand = idx.search_device(t["q_indptr"], t["q_indices"], torch.from_numpy(queries[2]).to(dev), 1000,
                                params=nsparse.SeismicSearchParameters(10, 0.9))
    t["cand"] = cand
    return docs, queries, cand.cpu().numpy(), t


def _scores64(D, Q, cand, rows=None):
    rows = range(cand.shape[0]) if rows is None else rows
    out = {}
    for q in rows:
        c = cand[q]
        s = np.zeros(len(c))
        ok = c >= 0
        s[ok] = (D[c[ok]] @ Q[q].T).toarray().ravel()
        out[q] = s
    return out


def test_c4_rescore_forward_backward_and_finite_differences(nsparse, c4):
    import torch

    from nsparse.rescore import info_nce, rescore

    docs, queries, cand, t = c4
    assert (cand[:, :10] >= 0).all()  # SEISMIC found at least 10 candidates per query
    Q = _csr(*queries)
    D = _csr(*docs)
    qv = torch.from_numpy(queries[2]).cuda().requires_grad_(True)
    dv = torch.from_numpy(docs[2]).cuda().requires_grad_(True)
    S = rescore(qv, dv, t["q_indptr"], t["q_indices"], t["cand"], t["d_indptr"], t["d_indices"], DIM)
    want = _scores64(D, Q, cand)
    Sn = S.detach().cpu().numpy()
    for q in range(cand.shape[0]):
        np.testing.assert_allclose(Sn[q], want[q], rtol=1e-5, atol=1e-6)

    # backward against the CPU reference gradient of sum(G * S)
    G = np.random.default_rng(0).standard_normal(Sn.shape)
    S.backward(torch.from_numpy(G.astype(np.float32)).cuda(), retain_graph=True)
    valid = cand >= 0
    gq = np.zeros(len(queries[2]))
    for q in range(cand.shape[0]):
        a, b = queries[0][q], queries[0][q + 1]
        c = cand[q][valid[q]]
        # d S[q, c] / d q_i = D[c, term_i]
        Dc = D[c][:, queries[1][a:b].astype(np.int64)].toarray()  # candidates x query entries
        gq[a:b] = G[q][valid[q]] @ Dc
    np.testing.assert_allclose(qv.grad.cpu().numpy(), gq, rtol=1e-4, atol=1e-4)
    import scipy.sparse as sp

    rr = np.repeat(np.arange(cand.shape[0]), cand.shape[1])[valid.ravel()]
    W = sp.csr_matrix((G.ravel()[valid.ravel()], (cand.ravel()[valid.ravel()], rr)),
                      shape=(len(docs[0]) - 1, cand.shape[0]))  # doc x query weights
    M = (W @ Q).tocsr()  # doc x term: sum over (q, c) of G[q, c] * Q[q, t]
    touched = np.unique(cand[valid])
    gd_want = np.zeros(len(docs[2]))
    for d in touched:
        a, b = docs[0][d], docs[0][d + 1]
        row = M[d]
        gd_want[a:b] = row[:, docs[1][a:b].astype(np.int64)].toarray().ravel()
    np.testing.assert_allclose(dv.grad.cpu().numpy(), gd_want, rtol=1e-4, atol=1e-4)

    # InfoNCE gradient against FP64 central differences (rows a coordinate touches)
    qv.grad = None
    dv.grad = None
    vt = torch.from_numpy(valid).cuda()
    info_nce(S, vt).backward()
    tau, eps, n_q = 0.05, 1e-3, cand.shape[0]

    def row_loss(s, q):
        z = np.where(valid[q], s / tau, -np.inf)
        m = z.max()
        return (m + np.log(np.exp(z - m).sum()) - z[0]) / n_q

    # the CSR data arrays keep the rows' entry order, so entry i / p is data[i] / data[p]
    rng = np.random.default_rng(1)
    for i in rng.choice(len(queries[2]), 5, replace=False):
        q = int(np.searchsorted(queries[0], i, side="right") - 1)
        fd = 0.0
        for sgn in (1, -1):
            Q.data[i] += sgn * eps
            fd += sgn * row_loss(_scores64(D, Q, cand, [q])[q], q)
            Q.data[i] -= sgn * eps
        fd /= 2 * eps
        assert abs(qv.grad[i].item() - fd) <= 1e-4 * max(abs(fd), 1e-2), (i, qv.grad[i].item(), fd)
    for d in rng.choice(cand[:, 0], 5, replace=False):
        p = int(docs[0][d]) + int(rng.integers(0, docs[0][d + 1] - docs[0][d]))
        rows = np.nonzero((cand == d).any(axis=1))[0]
        fd = 0.0
        for sgn in (1, -1):
            D.data[p] += sgn * eps
            sc = _scores64(D, Q, cand, rows)
            fd += sgn * sum(row_loss(sc[q], q) for q in rows)
            D.data[p] -= sgn * eps
        fd /= 2 * eps
        assert abs(dv.grad[p].item() - fd) <= 1e-4 * max(abs(fd), 1e-2), (p, dv.grad[p].item(), fd)
