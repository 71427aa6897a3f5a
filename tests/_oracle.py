"""ctypes access to the oracle (oracle/_build/liboracle.so) and, where it was
built, to the reference's own build-side code (oracle/_ref/libnsparse_ref.so).
Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libnsparse_ref.so")
P = ctypes.c_void_p
f32 = ctypes.c_float


def _ensure_built():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)


_ensure_built()
lib = ctypes.CDLL(ORACLE_SO)
lib.oracle_last_error.restype = ctypes.c_char_p
lib.oracle_sq_decode.restype = ctypes.c_float
lib.oracle_sq_decode.argtypes = [f32, f32, f32, f32, f32]
lib.oracle_mt19937_nth.restype = ctypes.c_uint32
lib.oracle_mt19937_nth.argtypes = [ctypes.c_uint32, ctypes.c_int]
lib.oracle_sq_decode16.restype = ctypes.c_float
lib.oracle_sq_decode16.argtypes = [f32, f32, f32, f32, f32]
ref = ctypes.CDLL(REF_SO) if os.path.exists(REF_SO) else None
if ref is not None:
    ref.ref_sq_decode.restype = ctypes.c_float
    ref.ref_sq_decode.argtypes = [f32, f32, f32, f32, f32]
    ref.ref_sq_decode16.restype = ctypes.c_float
    ref.ref_sq_decode16.argtypes = [f32, f32, f32, f32, f32]
    ref.ref_dot_u16.restype = ctypes.c_float
    ref.ref_dot_u8.restype = ctypes.c_float
    ref.ref_dot_f32.restype = ctypes.c_float

STAT_FIELDS = ["lists", "terms_found", "csc_entries", "clusters_tested", "clusters_opened", "docs_read",
               "docs_scored", "nnz_scored"]


def _p(a):
    return a.ctypes.data_as(P)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(lib.oracle_last_error().decode())


class OracleIndex:
    """The CPU restatement. sq: 0 FP32, 1 8-bit codes, 2 16-bit codes.
    tie_mode 0 = libstdc++ ties (the reference's), 1 = exact float ties broken
    by id (what the device path implements)."""

    def __init__(self, dim, sq, lam, beta, alpha, seed, vmin=0.0, vmax=1.0, tie_mode=1):
        self.h = P()
        self.sq = sq
        _chk(lib.oracle_create(dim, int(sq), f32(vmin), f32(vmax), lam, beta, f32(alpha), seed, tie_mode,
                               ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:
            lib.oracle_free(self.h)

    def add(self, ip, ix, v):
        ip = np.ascontiguousarray(ip, np.int32)
        ix = np.ascontiguousarray(ix, np.uint16)
        v = np.ascontiguousarray(v, np.float32)
        _chk(lib.oracle_add(self.h, len(ip) - 1, _p(ip), _p(ix), _p(v)))

    def build(self, threads=8):
        _chk(lib.oracle_build(self.h, threads))

    @classmethod
    def load(cls, blob, tie_mode=0):
        """An oracle index read from write_index bytes (the library's or the oracle's)."""
        self = cls.__new__(cls)
        self.h = P()
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        _chk(lib.oracle_load(buf, ctypes.c_size_t(len(blob)), tie_mode, ctypes.byref(self.h)))
        self.sq = bytes(blob[:4]) == b"SESQ"
        return self

    @classmethod
    def load_from_index(cls, nsparse, index, tie_mode=0):
        """Oracle copy of a library index: its write_index stream is handed over
        as a raw buffer (no Python bytes object, so no 2 GiB limit)."""
        import ctypes as ct

        buf = ct.POINTER(ct.c_uint8)()
        size = ct.c_size_t()
        nsparse._lib.check(nsparse._lib.lib.nsparse_write_index_buffer(index.handle, ct.byref(buf), ct.byref(size)))
        try:
            self = cls.__new__(cls)
            self.h = P()
            _chk(lib.oracle_load(ct.cast(buf, P), size, tie_mode, ct.byref(self.h)))
            self.sq = ct.string_at(buf, 4) == b"SESQ"
            return self
        finally:
            nsparse._lib.lib.nsparse_buffer_free(buf)

    def search_idmap(self, i2e, ip, ix, v, k, cut=10, hf=1.0, sq_range=None, selector=None, negate=False,
                     threads=8):
        """IDMapIndex::search with this index as the delegate: i2e[internal] =
        external id; the selector is on external ids; labels come back external."""
        ip = np.ascontiguousarray(ip, np.int32)
        ix = np.ascontiguousarray(ix, np.uint16)
        v = np.ascontiguousarray(v, np.float32)
        i2e = np.ascontiguousarray(i2e, np.int32)
        n = len(ip) - 1
        d = np.empty((n, k), np.float32)
        lab = np.empty((n, k), np.int32)
        kind, sn, ids, bm = -1, 0, None, None
        if selector is not None:
            if selector[0] in ("array", "set"):
                kind = 0 if selector[0] == "array" else 1
                ids = np.ascontiguousarray(selector[1], np.int32)
                sn = len(ids)
            else:
                kind = 3
                bm = np.ascontiguousarray(selector[1], np.uint32)
                sn = int(selector[2])
        qsq = 1 if sq_range is not None else 0
        lo, hi = sq_range if sq_range is not None else (0.0, 1.0)
        _chk(lib.oracle_search_idmap(self.h, _p(i2e), ctypes.c_int64(len(i2e)), n, _p(ip), _p(ix), _p(v), k, cut,
                                     f32(hf), qsq, f32(lo), f32(hi), kind, ctypes.c_int64(sn),
                                     _p(ids) if ids is not None else None, _p(bm) if bm is not None else None,
                                     int(negate), _p(d), _p(lab), threads))
        return d, lab

    def score_summaries(self, ip, ix, v, lists, cap=4096):
        """Summary scores of query q against list lists[q] (all its entries, SQ
        as codes): a list of float32 arrays, one per query."""
        ip = np.ascontiguousarray(ip, np.int32)
        ix = np.ascontiguousarray(ix, np.uint16)
        v = np.ascontiguousarray(v, np.float32)
        lists = np.ascontiguousarray(lists, np.int32)
        nq = len(ip) - 1
        out = np.zeros((nq, cap), np.float32)
        n = np.zeros(nq, np.int32)
        _chk(lib.oracle_score_summaries(self.h, nq, _p(ip), _p(ix), _p(v), _p(lists), cap, _p(out), _p(n)))
        return [out[q, : n[q]].copy() for q in range(nq)]

    def set_lists(self, lists):
        """Hand-made posting lists: lists[l] is a list of clusters, each a pair
        (doc ids, {term: summary value}); SQ summary values are codes."""
        nc, cdocs, docs, cterms, terms, vals = [], [], [], [], [], []
        for clusters in lists:
            nc.append(len(clusters))
            for ds, summ in clusters:
                cdocs.append(len(ds))
                docs.extend(ds)
                cterms.append(len(summ))
                for t in sorted(summ):
                    terms.append(t)
                    vals.append(summ[t])
        a = [np.ascontiguousarray(x, dt) for x, dt in ((nc, np.int32), (cdocs, np.int32), (docs, np.int32),
                                                        (cterms, np.int32), (terms, np.uint16), (vals, np.float32))]
        _chk(lib.oracle_set_lists(self.h, len(lists), *[_p(x) for x in a]))

    def copy_from(self, other):
        _chk(lib.oracle_copy_index(self.h, other.h))

    def to_library_index(self, nsparse):
        """The library's copy of this index: the write_index stream handed to
        nsparse_read_index_buffer as a raw buffer (no Python bytes object, so no
        2 GiB limit)."""
        buf = ctypes.POINTER(ctypes.c_uint8)()
        sz = ctypes.c_size_t()
        _chk(lib.oracle_serialize(self.h, ctypes.byref(buf), ctypes.byref(sz)))
        try:
            h = ctypes.c_void_p()
            nsparse._lib.check(nsparse._lib.lib.nsparse_read_index_buffer(ctypes.cast(buf, P), sz, 0, ctypes.byref(h)))
            return nsparse.Index(h.value)
        finally:
            lib.oracle_buffer_free(buf)

    def serialize(self):
        buf = ctypes.POINTER(ctypes.c_uint8)()
        sz = ctypes.c_size_t()
        _chk(lib.oracle_serialize(self.h, ctypes.byref(buf), ctypes.byref(sz)))
        out = ctypes.string_at(buf, sz.value)
        lib.oracle_buffer_free(buf)
        return out

    def search(self, ip, ix, v, k, cut=10, hf=1.0, sq_range=None, selector=None, negate=False, threads=8,
               with_stats=False):
        """selector: None, ("array", ids), ("set", ids) or ("bitmap", words, nbits)."""
        ip = np.ascontiguousarray(ip, np.int32)
        ix = np.ascontiguousarray(ix, np.uint16)
        v = np.ascontiguousarray(v, np.float32)
        n = len(ip) - 1
        d = np.empty((n, k), np.float32)
        lab = np.empty((n, k), np.int32)
        stats = np.zeros((n, 8), np.int64)
        kind, sn, ids, bm = -1, 0, None, None
        if selector is not None:
            if selector[0] in ("array", "set"):
                kind = 0 if selector[0] == "array" else 1
                ids = np.ascontiguousarray(selector[1], np.int32)
                sn = len(ids)
            else:
                kind = 3
                bm = np.ascontiguousarray(selector[1], np.uint32)
                sn = int(selector[2])
        qsq = 1 if sq_range is not None else 0
        lo, hi = sq_range if sq_range is not None else (0.0, 1.0)
        _chk(lib.oracle_search(self.h, n, _p(ip), _p(ix), _p(v), k, cut, f32(hf), qsq, f32(lo), f32(hi), kind,
                               ctypes.c_int64(sn), _p(ids) if ids is not None else None,
                               _p(bm) if bm is not None else None, int(negate), _p(d), _p(lab), _p(stats),
                               threads))
        return (d, lab, stats) if with_stats else (d, lab)


def search_mixed(o, queries, ks, k_max, cut, hf, sel_of_query=None, bitmaps=(), n_docs=0, threads=8):
    """The oracle's results for a batch with a k per query and, optionally, a
    bitmap selector per query (sel_of_query[i] indexes `bitmaps`, uint32 words
    over n_docs ids; -1 = none), laid out as the device's batched search writes
    them: row i holds ks[i] results, padded with (-1.0, -1) to k_max."""
    ip, ix, v = (np.asarray(a) for a in queries)
    n = len(ip) - 1
    ks = np.asarray(ks, np.int32)
    sel = np.full(n, -1, np.int32) if sel_of_query is None else np.asarray(sel_of_query, np.int32)
    D = np.full((n, k_max), -1.0, np.float32)
    L = np.full((n, k_max), -1, np.int32)
    for k, s in sorted({(int(a), int(b)) for a, b in zip(ks, sel)}):
        rows = np.nonzero((ks == k) & (sel == s))[0]
        lens = ip[rows + 1] - ip[rows]
        sip = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        take = np.concatenate([np.arange(ip[r], ip[r + 1]) for r in rows]) if len(rows) else np.zeros(0, int)
        selector = None if s < 0 else ("bitmap", bitmaps[s], n_docs)
        d, lab = o.search(sip, ix[take], v[take], k, cut=cut, hf=hf, selector=selector, threads=threads)
        D[rows, :k] = d
        L[rows, :k] = lab
    return D, L


def same_up_to_ties(d_ref, l_ref, d_got, l_got):
    """One result row against another, apart from exact score ties: the same
    scores in the same order, the same doc ids at every score above the k-th,
    and padding in the same places (docs tied with the k-th may be any of the
    tied ones, and equal scores may list their docs in any order)."""
    if not np.array_equal(d_ref, d_got):
        return False
    above = d_ref > d_ref[-1]
    if not np.array_equal(np.sort(l_ref[above]), np.sort(l_got[above])):
        return False
    return bool(np.array_equal(l_ref < 0, l_got < 0))


def ref_build_bytes(dim, sq, ip, ix, v, lam, beta, alpha, seed, vmin=0.0, vmax=1.0):
    """write_index bytes produced by the reference's own build-side code."""
    ip = np.ascontiguousarray(ip, np.int32)
    ix = np.ascontiguousarray(ix, np.uint16)
    v = np.ascontiguousarray(v, np.float32)
    buf = ctypes.POINTER(ctypes.c_uint8)()
    sz = ctypes.c_size_t()
    rc = ref.ref_build_serialize(dim, int(sq), f32(vmin), f32(vmax), len(ip) - 1, _p(ip), _p(ix), _p(v), lam,
                                 beta, f32(alpha), ctypes.c_uint32(seed), ctypes.byref(buf), ctypes.byref(sz))
    if rc != 0:
        raise RuntimeError(ref.ref_last_error().decode())
    out = ctypes.string_at(buf, sz.value)
    ref.ref_buffer_free(buf)
    return out


def ref_score_summaries(dim, sq, docs, lam, beta, alpha, seed, queries, lists, vmin=0.0, vmax=1.0, cap=4096):
    """Summary scores from the reference's own InvertedListClusters::
    score_summaries_transposed over its own build (see OracleIndex.score_summaries)."""
    ip, ix, v = (np.ascontiguousarray(a, dt) for a, dt in zip(docs, (np.int32, np.uint16, np.float32)))
    qip, qix, qv = (np.ascontiguousarray(a, dt) for a, dt in zip(queries, (np.int32, np.uint16, np.float32)))
    lists = np.ascontiguousarray(lists, np.int32)
    nq = len(qip) - 1
    out = np.zeros((nq, cap), np.float32)
    n = np.zeros(nq, np.int32)
    rc = ref.ref_score_summaries(dim, int(sq), f32(vmin), f32(vmax), len(ip) - 1, _p(ip), _p(ix), _p(v), lam, beta,
                                 f32(alpha), ctypes.c_uint32(seed), nq, _p(qip), _p(qix), _p(qv), _p(lists), cap,
                                 _p(out), _p(n))
    if rc != 0:
        raise RuntimeError(ref.ref_last_error().decode())
    return [out[q, : n[q]].copy() for q in range(nq)]


def cluster_order(keys):
    keys = np.ascontiguousarray(keys, np.float32)
    out = np.empty(len(keys), np.int64)
    _chk(lib.oracle_cluster_order(_p(keys), len(keys), _p(out)))
    return out


def top_k_tokens_f32(terms, weights, k):
    terms = np.ascontiguousarray(terms, np.uint16)
    weights = np.ascontiguousarray(weights, np.float32)
    out = np.empty(len(terms) + 1, np.uint16)
    n = ctypes.c_int()
    _chk(lib.oracle_top_k_tokens_f32(_p(terms), _p(weights), len(terms), k, _p(out), ctypes.byref(n)))
    return out[: n.value]


def brute_force_top_k(docs, queries, dim, k):
    """Exact top-k by dot product, score desc then id asc, positive scores only
    (the reference's python_tests/oracle.py:28-47), in float64 via scipy."""
    import scipy.sparse as sp

    D = sp.csr_matrix((docs[2].astype(np.float64), docs[1].astype(np.int64), docs[0].astype(np.int64)),
                      shape=(len(docs[0]) - 1, dim))
    Q = sp.csr_matrix((queries[2].astype(np.float64), queries[1].astype(np.int64),
                       queries[0].astype(np.int64)), shape=(len(queries[0]) - 1, dim))
    S = (Q @ D.T).toarray()
    labels = np.full((S.shape[0], k), -1, np.int64)
    for q in range(S.shape[0]):
        (el,) = np.nonzero(S[q] > 0.0)
        order = el[np.lexsort((el, -S[q][el]))][:k]
        labels[q, : len(order)] = order
    return labels


def recall_at_k(got, want):
    per = []
    for g, w in zip(got, want):
        ws = {int(i) for i in w if i >= 0}
        if not ws:
            continue
        per.append(len({int(i) for i in g if i >= 0} & ws) / len(ws))
    return float(np.mean(per)) if per else 1.0


class OracleInverted:
    """The CPU restatement of InvertedIndex (MaxScore over 4096-doc windows).
    tie_mode as for OracleIndex."""

    def __init__(self, dim, tie_mode=1):
        self.h = P()
        _chk(lib.oracle_inv_create(dim, tie_mode, ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.oracle_inv_free(self.h)
            self.h = None

    def add(self, ip, ix, v):
        ip = np.ascontiguousarray(ip, np.int32)
        ix = np.ascontiguousarray(ix, np.uint16)
        v = np.ascontiguousarray(v, np.float32)
        _chk(lib.oracle_inv_add(self.h, len(ip) - 1, _p(ip), _p(ix), _p(v)))

    def build(self):
        _chk(lib.oracle_inv_build(self.h))

    def num_vectors(self):
        lib.oracle_inv_num_vectors.restype = ctypes.c_int64
        return int(lib.oracle_inv_num_vectors(self.h))

    def search(self, ip, ix, v, k, selector=None, negate=False, threads=8, fill=(-1.0, -1)):
        """selector: None, ("array", ids), ("set", ids) or ("bitmap", words, nbits).
        Outputs start as `fill`; an unbuilt index leaves them so."""
        ip = np.ascontiguousarray(ip, np.int32)
        ix = np.ascontiguousarray(ix, np.uint16)
        v = np.ascontiguousarray(v, np.float32)
        n = len(ip) - 1
        d = np.full((n, k), fill[0], np.float32)
        lab = np.full((n, k), fill[1], np.int32)
        kind, sn, ids, bm = -1, 0, None, None
        if selector is not None:
            if selector[0] in ("array", "set"):
                kind = 0 if selector[0] == "array" else 1
                ids = np.ascontiguousarray(selector[1], np.int32)
                sn = len(ids)
            else:
                kind = 3
                bm = np.ascontiguousarray(selector[1], np.uint32)
                sn = int(selector[2])
        _chk(lib.oracle_inv_search(self.h, n, _p(ip), _p(ix), _p(v), k, kind, ctypes.c_int64(sn),
                                   _p(ids) if ids is not None else None, _p(bm) if bm is not None else None,
                                   int(negate), _p(d), _p(lab), threads))
        return d, lab

    def serialize(self):
        buf = ctypes.POINTER(ctypes.c_uint8)()
        sz = ctypes.c_size_t()
        _chk(lib.oracle_inv_serialize(self.h, ctypes.byref(buf), ctypes.byref(sz)))
        out = ctypes.string_at(buf, sz.value)
        lib.oracle_buffer_free(buf)
        return out


class TermMajor:
    """The corpus grouped by term (stable: doc ids ascend inside a term), for
    exact FP32 scores summed over a query's entries in query order: for each
    entry, every posting of its term adds weight * value (float32 product, then
    a float32 add) -- the order of the device's term-at-a-time scorer. Rows
    must hold each term at most once."""

    def __init__(self, docs, dim):
        ip, ix, v = docs
        self.n = len(ip) - 1
        order = np.argsort(ix, kind="stable")
        self.rows = np.repeat(np.arange(self.n), np.diff(ip))[order]
        self.vals = v[order]
        self.ptr = np.zeros(dim + 1, np.int64)
        np.cumsum(np.bincount(ix, minlength=dim), out=self.ptr[1:])

    def scores(self, q_idx, q_val):
        acc = np.zeros(self.n, np.float32)
        for t, w in zip(q_idx, q_val):
            a, b = self.ptr[t], self.ptr[t + 1]
            r = self.rows[a:b]
            acc[r] = acc[r] + np.float32(w) * self.vals[a:b]
        return acc

    def top_k(self, queries, k, allowed=None):
        """Exact top-k of positive scores (score desc, doc id asc) in that order."""
        qip, qix, qv = queries
        nq = len(qip) - 1
        labels = np.full((nq, k), -1, np.int32)
        dists = np.full((nq, k), -1.0, np.float32)
        for q in range(nq):
            s = self.scores(qix[qip[q]:qip[q + 1]], qv[qip[q]:qip[q + 1]])
            keep = s > 0
            if allowed is not None:
                keep &= allowed
            (el,) = np.nonzero(keep)
            order = el[np.lexsort((el, -s[el]))][:k]
            labels[q, :len(order)] = order
            dists[q, :len(order)] = s[order]
        return labels, dists


def ref_inverted_bytes(dim, ip, ix, v):
    """INVT bytes of an index built by the reference's own posting-list code."""
    ip = np.ascontiguousarray(ip, np.int32)
    ix = np.ascontiguousarray(ix, np.uint16)
    v = np.ascontiguousarray(v, np.float32)
    buf = ctypes.POINTER(ctypes.c_uint8)()
    sz = ctypes.c_size_t()
    rc = ref.ref_inverted_serialize(dim, len(ip) - 1, _p(ip), _p(ix), _p(v), ctypes.byref(buf), ctypes.byref(sz))
    if rc != 0:
        raise RuntimeError(ref.ref_last_error().decode())
    out = ctypes.string_at(buf, sz.value)
    ref.ref_buffer_free(buf)
    return out


def exact_scores64(docs, queries, dim):
    """Every query's score against every doc in float64 (scipy), n_q x n_docs."""
    import scipy.sparse as sp

    D = sp.csr_matrix((docs[2].astype(np.float64), docs[1].astype(np.int64), docs[0].astype(np.int64)),
                      shape=(len(docs[0]) - 1, dim))
    Q = sp.csr_matrix((queries[2].astype(np.float64), queries[1].astype(np.int64),
                       queries[0].astype(np.int64)), shape=(len(queries[0]) - 1, dim))
    return (Q @ D.T).toarray()


def assert_topk_matches(labels, dists, scores64, k, allowed=None, rtol=1e-5):
    """labels/dists against the exact top-k of positive scores (score desc, id
    asc) from float64 scores: scores agree within rtol, and a label may differ
    from the exact ranking only where its exact score ties the expected one
    within rtol (a float32 sum can order such near-ties either way)."""
    for q in range(scores64.shape[0]):
        s = scores64[q].copy()
        if allowed is not None:
            s[~allowed] = 0.0
        (el,) = np.nonzero(s > 0.0)
        order = el[np.lexsort((el, -s[el]))][:k]
        want_l = np.full(k, -1, np.int64)
        want_l[: len(order)] = order
        got_l = labels[q].astype(np.int64)
        assert np.array_equal(got_l < 0, want_l < 0), (q, got_l, want_l)
        m = want_l >= 0
        np.testing.assert_allclose(dists[q][m], s[want_l[m]], rtol=rtol, atol=1e-6)
        assert np.all(dists[q][~m] == -1.0)
        for i in np.nonzero(got_l != want_l)[0]:
            g, w = got_l[i], want_l[i]
            assert g >= 0 and abs(s[g] - s[w]) <= rtol * abs(s[w]) + 1e-6, (q, i, g, w, s[g], s[w])
