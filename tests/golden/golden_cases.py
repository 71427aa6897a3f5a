"""The cases behind tests/golden/ref_fixtures.json: seeded corpora (numpy PCG64,
stable across numpy releases) and the parameters each was built with."""
import hashlib

import numpy as np


def corpus(n, dim, nnz, seed, zipf=1.1, negative=False, min_len=2):
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, dim + 1) ** zipf
    p /= p.sum()
    ip, ix, v = [0], [], []
    for _ in range(n):
        m = int(np.clip(round(np.exp(np.log(nnz) + 0.35 * rng.standard_normal())), min_len, dim))
        t = np.sort(rng.choice(dim, size=m, replace=False, p=p))
        w = rng.lognormal(0, 0.7, m).astype(np.float32)
        if negative:
            w[rng.random(m) < 0.2] *= -1
        ix.append(t)
        v.append(w)
        ip.append(ip[-1] + m)
    return np.array(ip, np.int32), np.concatenate(ix).astype(np.uint16), np.concatenate(v).astype(np.float32)


def _resolve(lam, beta, n):
    # seismic_common.h:130-144, as the reference's driver resolves -1
    lam_r = max(int(np.float32(0.0005) * np.float32(n)), 160) if lam == -1 else lam
    beta_r = int(np.float32(lam_r) * np.float32(0.1)) if beta == -1 else beta
    return lam_r, beta_r


BUILD_CASES = []
for _sq in (0, 1, 2):  # FP32, 8-bit codes, 16-bit codes
    for _n, _dim, _nnz, _lam, _beta in [(2000, 1024, 30, 100, 10), (3000, 4096, 60, 300, 30), (500, 256, 20, -1, -1),
                                        (400, 300, 12, 0, 0)]:
        _lr, _br = _resolve(_lam, _beta, _n)
        BUILD_CASES.append({"sq": _sq, "n": _n, "dim": _dim, "nnz": _nnz, "lam": _lam, "beta": _beta,
                            "lam_resolved": _lr, "beta_resolved": _br, "alpha": 0.4, "seed": 7,
                            "corpus_seed": _n + _dim})

INVERTED_CASES = [{"n": 300, "dim": 64, "nnz": 8, "corpus_seed": 1, "negative": False},
                  {"n": 5000, "dim": 512, "nnz": 20, "corpus_seed": 2, "negative": False},
                  {"n": 800, "dim": 128, "nnz": 6, "corpus_seed": 3, "negative": True}]

SQ_RANGES = [(0.0, 3.0), (0.01, 43.9), (-2.0, 5.0)]


def build_corpus(case):
    return corpus(case["n"], case["dim"], case["nnz"], seed=case["corpus_seed"])


def inverted_corpus(case):
    ip, ix, v = corpus(case["n"], case["dim"], case["nnz"], case["corpus_seed"], negative=case["negative"],
                       min_len=1)
    v[::17] = 0.0  # zero postings stay in the lists; maxima start from 0
    return ip, ix, v


def sq_values():
    rng = np.random.default_rng(0)
    return np.concatenate([rng.lognormal(0, 1.5, 20000), [0.0, -1.0, 1e9, 3.0, 0.5, 2.5],
                           np.float32(np.arange(0, 256) / 255.0 * 3.0)]).astype(np.float32)


def summary_pairs(case, n_queries=12, lists_per_query=5):
    """(query, posting list) pairs for summary scoring on a build case: seeded
    queries over the case's vocabulary; each query is paired with the lists of
    its lists_per_query largest-weight terms and of its first term. Query 0
    repeats its first term at the end with another weight (a repeated entry
    adds again in score_summaries). Returns the queries one row per pair (CSR)
    and the list of each row."""
    qip, qix, qv = corpus(n_queries, case["dim"], max(4, case["nnz"] // 2), seed=case["corpus_seed"] + 1)
    rows, lists = [], []
    for q in range(n_queries):
        t = qix[qip[q]:qip[q + 1]].copy()
        w = qv[qip[q]:qip[q + 1]].copy()
        if q == 0:
            t = np.append(t, t[0]).astype(np.uint16)
            w = np.append(w, np.float32(2.5) * w[0]).astype(np.float32)
        picks = list(t[np.argsort(-w, kind="stable")[:lists_per_query]]) + [t[0]]
        for term in picks:
            rows.append((t, w))
            lists.append(int(term))
    ip = np.zeros(len(rows) + 1, np.int32)
    ip[1:] = np.cumsum([len(r[0]) for r in rows])
    ix = np.concatenate([r[0] for r in rows]).astype(np.uint16)
    v = np.concatenate([r[1] for r in rows]).astype(np.float32)
    return ip, ix, v, np.array(lists, np.int32)


def summary_digest(scores):
    """SHA-256 over every pair's cluster count (int32) and scores (float32 bits)."""
    h = hashlib.sha256()
    for sc in scores:
        h.update(np.int32(len(sc)).tobytes())
        h.update(np.ascontiguousarray(sc, np.float32).tobytes())
    return h.hexdigest()
