"""Differentiable re-scoring (nsparse.rescore): forward scores against an
exact CPU computation, gradients against a CPU reference gradient and FP64
central finite differences (eps 1e-3, relative error <= 1e-4), on candidates
mined by the GPU SEISMIC search, with an InfoNCE loss (temperature 0.05, first
candidate positive). A query term given twice takes the value of its last
entry (the dense query of compute_similarity, seismic_common.h:78-103, built
by SparseVectors::get_dense_vector): its earlier entries get gradient 0."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIM = 30522


@pytest.fixture(scope="module")
def setup(nsparse):
    import torch

    from nsparse import synth

    docs = synth.make_corpus(synth.doc_spec(seed=40), 5000)
    queries = synth.make_corpus(synth.query_spec(seed=41), 16)
    idx = nsparse.index_factory(DIM, "seismic,lambda=300|beta=30|alpha=0.4|seed=0")
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    _, cand = idx.search(len(queries[0]) - 1, *queries, 50, nsparse.SeismicSearchParameters(10, 0.9))
    dev = torch.device("cuda")
    t = dict(
        q_indptr=torch.from_numpy(queries[0]).to(dev), q_indices=torch.from_numpy(queries[1].view(np.int16)).to(dev),
        d_indptr=torch.from_numpy(docs[0]).to(dev), d_indices=torch.from_numpy(docs[1].view(np.int16)).to(dev),
        cand=torch.from_numpy(cand).to(dev))
    return docs, queries, cand, t


def dense_query(qix, qv, lo, hi):
    """term -> (value, entry index) of the last entry of each term."""
    qmap = {}
    for i in range(lo, hi):
        qmap[int(qix[i])] = (qv[i], i)
    return qmap


def cpu_scores(docs, queries, cand, qv, dv):
    """float64 exact sparse dot products against the dense query."""
    dip, dix, _ = docs
    qip, qix, _ = queries
    n_q, n_c = cand.shape
    S = np.zeros((n_q, n_c))
    for q in range(n_q):
        qmap = dense_query(qix, qv, qip[q], qip[q + 1])
        for c in range(n_c):
            d = cand[q, c]
            if d < 0:
                continue
            S[q, c] = sum(dv[p] * qmap.get(int(dix[p]), (0.0, -1))[0] for p in range(dip[d], dip[d + 1]))
    return S


def cpu_grads(docs, queries, cand, qv, dv, G):
    dip, dix, _ = docs
    qip, qix, _ = queries
    gq = np.zeros(len(qv))
    gd = np.zeros(len(dv))
    for q in range(cand.shape[0]):
        qmap = dense_query(qix, qv, qip[q], qip[q + 1])
        last = {i for _, i in qmap.values()}
        for c in range(cand.shape[1]):
            d = cand[q, c]
            if d < 0:
                continue
            for i in range(qip[q], qip[q + 1]):
                if i in last:
                    gq[i] += G[q, c] * sum(dv[p] for p in range(dip[d], dip[d + 1]) if int(dix[p]) == int(qix[i]))
            for p in range(dip[d], dip[d + 1]):
                gd[p] += G[q, c] * qmap.get(int(dix[p]), (0.0, -1))[0]
    return gq, gd


def test_forward_and_cpu_gradient(nsparse, setup):
    import torch

    from nsparse.rescore import rescore

    docs, queries, cand, t = setup
    qv = torch.from_numpy(queries[2]).cuda().requires_grad_(True)
    dv = torch.from_numpy(docs[2]).cuda().requires_grad_(True)
    S = rescore(qv, dv, t["q_indptr"], t["q_indices"], t["cand"], t["d_indptr"], t["d_indices"], DIM)
    want = cpu_scores(docs, queries, cand, queries[2].astype(np.float64), docs[2].astype(np.float64))
    np.testing.assert_allclose(S.detach().cpu().numpy(), want, rtol=1e-5, atol=1e-6)
    G = np.random.default_rng(0).standard_normal(S.shape).astype(np.float32)
    S.backward(torch.from_numpy(G).cuda())
    gq, gd = cpu_grads(docs, queries, cand, queries[2].astype(np.float64), docs[2].astype(np.float64),
                       G.astype(np.float64))
    np.testing.assert_allclose(qv.grad.cpu().numpy(), gq, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(dv.grad.cpu().numpy(), gd, rtol=1e-4, atol=1e-5)


def test_info_nce_gradient_matches_finite_differences(nsparse, setup):
    import torch

    from nsparse.rescore import info_nce, rescore

    docs, queries, cand, t = setup
    valid = t["cand"] >= 0
    qv = torch.from_numpy(queries[2]).cuda().requires_grad_(True)
    dv = torch.from_numpy(docs[2]).cuda().requires_grad_(True)
    loss = info_nce(rescore(qv, dv, t["q_indptr"], t["q_indices"], t["cand"], t["d_indptr"], t["d_indices"], DIM),
                    valid)
    loss.backward()

    # float64 loss on the CPU for central differences
    valid_np = cand >= 0

    def loss64(qv64, dv64):
        S = cpu_scores(docs, queries, cand, qv64, dv64) / 0.05
        S = np.where(valid_np, S, -np.inf)
        m = S.max(axis=1, keepdims=True)
        lse = m[:, 0] + np.log(np.exp(S - m).sum(axis=1))
        return float(np.mean(lse - S[:, 0]))

    qv64 = queries[2].astype(np.float64)
    dv64 = docs[2].astype(np.float64)
    rng = np.random.default_rng(1)
    eps = 1e-3
    for i in rng.choice(len(qv64), 6, replace=False):
        a, b = qv64.copy(), qv64.copy()
        a[i] += eps
        b[i] -= eps
        fd = (loss64(a, dv64) - loss64(b, dv64)) / (2 * eps)
        assert abs(qv.grad[i].item() - fd) <= 1e-4 * max(abs(fd), 1e-2), (i, qv.grad[i].item(), fd)
    touched = np.unique(np.concatenate([np.arange(docs[0][d], docs[0][d + 1]) for d in cand[:, :3].ravel()
                                        if d >= 0]))
    for p in rng.choice(touched, 6, replace=False):
        a, b = dv64.copy(), dv64.copy()
        a[p] += eps
        b[p] -= eps
        fd = (loss64(qv64, a) - loss64(qv64, b)) / (2 * eps)
        assert abs(dv.grad[p].item() - fd) <= 1e-4 * max(abs(fd), 1e-2), (p, dv.grad[p].item(), fd)


def test_two_stage_composition(nsparse):
    """The two-stage search as composed from the public API: an SQ search plan
    mines 20 candidates, `rescore` scores them in FP32, `torch.topk` keeps 10.
    The result is each query's exact FP32 top-10 among its candidates
    (float64 CPU scores; ties in FP32 accepted by score), and its recall
    against exact FP32 brute force is not below the SQ search's own top-10
    (a k=20 traversal prunes differently from a k=10 one, so its candidates
    need not contain that top-10: 0.05 of slack)."""
    import torch

    from nsparse import synth
    from nsparse.rescore import rescore

    print("two-stage composition", flush=True)
    docs = synth.make_corpus(synth.doc_spec(seed=70), 6000)
    queries = synth.make_corpus(synth.query_spec(seed=71), 24)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    idx = nsparse.index_factory(
        DIM, f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda=300|beta=30|alpha=0.4|seed=0")
    idx.add(len(docs[0]) - 1, *docs)
    idx.build()
    dev = torch.device("cuda")
    q_ip = torch.from_numpy(queries[0]).to(dev)
    q_ix = torch.from_numpy(queries[1].view(np.int16)).to(dev)
    q_v = torch.from_numpy(queries[2]).to(dev)
    d_ip = torch.from_numpy(docs[0]).to(dev)
    d_ix = torch.from_numpy(docs[1].view(np.int16)).to(dev)
    d_v = torch.from_numpy(docs[2]).to(dev)
    plan = idx.search_plan(q_ip, q_ix, q_v, 20, params=nsparse.SeismicSearchParameters(8, 0.8))
    _, cand = plan.launch()
    with torch.no_grad():
        s = rescore(q_v, d_v, q_ip, q_ix, cand, d_ip, d_ix, DIM)
    top = torch.topk(s.masked_fill(cand < 0, float("-inf")), 10, dim=1)
    labels = torch.gather(cand, 1, top.indices).cpu().numpy()
    torch.cuda.synchronize()
    c = cand.cpu().numpy()
    exact = cpu_scores(docs, queries, c, queries[2].astype(np.float64), docs[2].astype(np.float64))
    for q in range(c.shape[0]):
        valid = c[q] >= 0
        want = np.sort(exact[q][valid])[::-1][:10]
        got = np.array([exact[q][list(c[q]).index(d)] for d in labels[q] if d >= 0])
        np.testing.assert_allclose(got, want[:len(got)], rtol=1e-5)
    # recall against exact FP32 brute force over all docs, against the SQ search's own top-10
    _, sq_lab = idx.search(len(queries[0]) - 1, *queries, 10, nsparse.SeismicSearchParameters(8, 0.8))
    brute = nsparse.index_factory(DIM, "brutal")
    brute.add(len(docs[0]) - 1, *docs)
    _, truth = brute.search(len(queries[0]) - 1, *queries, 10)
    rec = lambda lab: np.mean([len(set(lab[i]) & set(truth[i])) / 10 for i in range(len(truth))])  # noqa: E731
    print(f"recall@10 two-stage {rec(labels):.3f}, SQ {rec(sq_lab):.3f}", flush=True)
    assert rec(labels) >= rec(sq_lab) - 0.05


def test_repeated_query_terms_take_the_last_entry(nsparse):
    """Queries that repeat terms (with different values) score as the
    brute-force index scores them (the last entry of a term wins), within
    1e-5 relative; the gradient of an earlier entry of a repeated term is 0 and
    that of the last entry is the sum over candidates of g * d_t."""
    import torch

    from nsparse import synth
    from nsparse.rescore import rescore

    docs = synth.make_corpus(synth.doc_spec(seed=44), 3000)
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=45), 12)
    rng = np.random.default_rng(46)
    # every query gets two extra entries repeating its first two terms with new values
    ip2, ix2, v2 = [0], [], []
    for q in range(len(qip) - 1):
        t = list(qix[qip[q]:qip[q + 1]])
        w = list(qv[qip[q]:qip[q + 1]])
        t += t[:2]
        w += list(rng.uniform(0.5, 3.0, 2).astype(np.float32))
        ix2 += t
        v2 += w
        ip2.append(len(ix2))
    queries = (np.array(ip2, np.int32), np.array(ix2, np.uint16), np.array(v2, np.float32))
    brute = nsparse.index_factory(DIM, "brutal")
    brute.add(len(docs[0]) - 1, *docs)
    bd, cand = brute.search(len(ip2) - 1, *queries, 30)
    dev = torch.device("cuda")
    qv_t = torch.from_numpy(queries[2]).to(dev).requires_grad_(True)
    dv_t = torch.from_numpy(docs[2]).to(dev)
    S = rescore(qv_t, dv_t, torch.from_numpy(queries[0]).to(dev), torch.from_numpy(queries[1].view(np.int16)).to(dev),
                torch.from_numpy(cand).to(dev), torch.from_numpy(docs[0]).to(dev),
                torch.from_numpy(docs[1].view(np.int16)).to(dev), DIM)
    valid = cand >= 0
    np.testing.assert_allclose(S.detach().cpu().numpy()[valid], bd[valid], rtol=1e-5, atol=1e-6)
    G = rng.standard_normal(S.shape).astype(np.float32)
    S.backward(torch.from_numpy(G).to(dev))
    gq, _ = cpu_grads(docs, queries, cand, queries[2].astype(np.float64), docs[2].astype(np.float64),
                      G.astype(np.float64))
    got = qv_t.grad.cpu().numpy()
    np.testing.assert_allclose(got, gq, rtol=1e-4, atol=1e-5)
    for q in range(len(ip2) - 1):
        assert got[ip2[q]] == 0.0 and got[ip2[q] + 1] == 0.0  # shadowed by the repeats at the end
