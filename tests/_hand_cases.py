"""Hand-made SEISMIC indexes whose search results are derived by hand from
the reference's traversal (seismic_index.cpp:53-118, 231-274;
seismic_scalar_quantized_index.cpp:67-135, 307-357; seismic_common.h:105-115;
utils/vector_process.h:18-30; utils/ranker.h:34-138), each built so that one
plausible mistake changes the expected labels. Test infrastructure only.

A case is an FP32 index over `dim` terms: `docs` are the forward-index rows
({term: weight}); `lists[l]` is posting list l, a sequence of clusters
(doc ids in stored order, {term: summary weight}); `query` is one row of
(term, weight) in CSR order. `want[tie_mode]` = (labels, distances) at k.
The derivation of each expectation is written beside it.
"""

CASES = {
    # cut >= q_len keeps the query's terms in CSR order: cuts [0, 1].
    # List 0 (first): one cluster, summary 2*1 = 2 -> d0 scores 2, heap {d0: 2}.
    # List 1 (not first, clusters in stored order): C0 summary 1*1 = 1 < 2
    # fails -> continue; C1 summary 1*5 = 5 >= 2 -> d2 scores 5 -> heap {d2: 5}.
    # A `break` on a later list would stop at C0 and return d0 (2.0).
    "later_list_continue": dict(
        dim=16,
        docs=[{0: 1.0}, {1: 1.0}, {1: 5.0}],
        lists={0: [([0], {0: 1.0})], 1: [([1], {1: 1.0}), ([2], {1: 5.0})]},
        query=[(0, 2.0), (1, 1.0)], cut=2, k=1, hf=1.0,
        want={0: ([2], [5.0]), 1: ([2], [5.0])}),
    # cut 1 < q_len: top-1 by weight is term 0 (1.0 > -5.0). Summary scores
    # use every query entry: C0 = 1*4 = 4, C1 = 1*1 + (-5)*2 = -9, C2 = 1*2 = 2.
    # First list, sorted descending: C0, C2, C1. heap_factor -1: C0 is taken
    # (heap empty), d0 scores 4; C2: 2 * -1 = -2 < 4 fails -> break. Continuing
    # instead would reach C1 (-9 * -1 = 9 >= 4) and take d1 (10 - 5 = 5).
    "first_list_break": dict(
        dim=16,
        docs=[{0: 4.0}, {0: 10.0, 1: 1.0}, {0: 2.0}],
        lists={0: [([0], {0: 4.0}), ([1], {0: 1.0, 1: 2.0}), ([2], {0: 2.0})]},
        query=[(0, 1.0), (1, -5.0)], cut=1, k=1, hf=-1.0,
        want={0: ([0], [4.0]), 1: ([0], [4.0])}),
    # cut 1 -> cuts [0] (3 > 2). Summaries over all query entries: C0 = 3*1 = 3,
    # C1 = 2*5 = 10 (from term 1, which is not a cut term). Sorted: C1, C0.
    # d1 scores 2*5 = 10; C0: 3 < 10 -> break. Scoring summaries over the cut
    # terms only would give C1 = 0, visit C0 first (d0 = 3) and stop: d0 (3.0).
    "summary_from_non_cut_term": dict(
        dim=16,
        docs=[{0: 1.0}, {1: 5.0}],
        lists={0: [([0], {0: 1.0}), ([1], {1: 5.0})]},
        query=[(0, 3.0), (1, 2.0)], cut=1, k=1, hf=1.0,
        want={0: ([1], [10.0]), 1: ([1], [10.0])}),
    # cuts [0, 1]; list 0 exists but is empty, which still ends the first list
    # (the loop sets first_list = false after every list below n_lists). List
    # 1 is then visited in stored order: C0 (summary 2) -> d0 scores 2*20 = 40;
    # C1 (10) and C2 (3) fail against 40 and are skipped. Had the empty list
    # kept first_list, list 1 would be sorted (C1, C2, C0): d1 = 10, then C2
    # 3 < 10 breaks: d1 (10.0).
    "empty_list_ends_first_list": dict(
        dim=16,
        docs=[{1: 20.0}, {1: 5.0}, {1: 1.5}],
        lists={0: [], 1: [([0], {1: 1.0}), ([1], {1: 5.0}), ([2], {1: 1.5})]},
        query=[(0, 3.0), (1, 2.0)], cut=2, k=1, hf=1.0,
        want={0: ([0], [40.0]), 1: ([0], [40.0])}),
    # Same docs and list 1, but the index has 8 lists and the first cut term
    # (9, CSR order) is >= n_lists: it is skipped without ending the first
    # list, so list 1 is the first list: sorted C1 (10), C2 (3), C0 (2); d1
    # scores 2*5 = 10; C2: 3 < 10 -> break. Treating the skipped term like an
    # empty list would visit list 1 in stored order: d0 (40.0).
    "out_of_range_term_keeps_first_list": dict(
        dim=16, n_lists=8,
        docs=[{1: 20.0}, {1: 5.0}, {1: 1.5}],
        lists={1: [([0], {1: 1.0}), ([1], {1: 5.0}), ([2], {1: 1.5})]},
        query=[(9, 3.0), (1, 2.0)], cut=2, k=1, hf=1.0,
        want={0: ([1], [10.0]), 1: ([1], [10.0])}),
    # Two clusters tie at summary 5. std::ranges::sort of <= 16 elements is a
    # plain insertion sort, which moves an element left only past strictly
    # smaller keys: the tie keeps stored order C0, C1. d0 scores 6, then C1:
    # 5 < 6 -> break. Ties taken in reverse (or an unstable partition) would
    # visit d1 first: d1 (7.0).
    "first_list_order_tie": dict(
        dim=16,
        docs=[{0: 6.0}, {0: 7.0}],
        lists={0: [([0], {0: 5.0}), ([1], {0: 5.0})]},
        query=[(0, 1.0)], cut=1, k=1, hf=1.0,
        want={0: ([0], [6.0]), 1: ([0], [6.0])}),
    # d2 and d1 both score 3 and are visited in that order (one cluster).
    # TopKHolder rejects a score <= the k-th once full: the libstdc++ tie mode
    # keeps d2. The id tie mode keeps the lower doc id, d1.
    "result_tie": dict(
        dim=16,
        docs=[{0: 1.0}, {0: 3.0}, {0: 3.0}],
        lists={0: [([2, 1], {0: 3.0})]},
        query=[(0, 1.0)], cut=1, k=1, hf=1.0,
        want={0: ([2], [3.0]), 1: ([1], [3.0])}),
    # A repeated query term: the cut (top-1 of weights 1, 3) is term 0 once.
    # Summary scoring adds each entry (C0 = 1*1 + 3*1 = 4, C1 = 1*2 + 3*2 = 8
    # -> sorted C1, C0), while the dense query keeps the last entry (term 0 ->
    # 3). k = 3: C1 -> d1 scores 3*2 = 6; C0 is visited too (the heap holds 1
    # of 3, not full): d0 scores 3*1 = 3, d2 3*0.5 = 1.5. Results d1, d0, d2 =
    # 6, 3, 1.5. Summing the repeated entries in the dense query instead would
    # give 8, 4, 2.
    "repeated_query_term": dict(
        dim=16,
        docs=[{0: 1.0}, {0: 2.0}, {0: 0.5}],
        lists={0: [([0, 2], {0: 1.0}), ([1], {0: 2.0})]},
        query=[(0, 1.0), (0, 3.0)], cut=1, k=3, hf=0.5,
        want={0: ([1, 0, 2], [6.0, 3.0, 1.5]), 1: ([1, 0, 2], [6.0, 3.0, 1.5])}),
}


def build_oracle(case, tie_mode):
    """The case as an oracle index (FP32, hand-made lists)."""
    import numpy as np

    from _oracle import OracleIndex

    dim = case["dim"]
    o = OracleIndex(dim, 0, 10, 5, 0.5, 0, tie_mode=tie_mode)
    ip, ix, v = [0], [], []
    for d in case["docs"]:
        for t in sorted(d):
            ix.append(t)
            v.append(d[t])
        ip.append(len(ix))
    o.add(np.array(ip, np.int32), np.array(ix, np.uint16), np.array(v, np.float32))
    n_lists = case.get("n_lists", dim)
    o.set_lists([case["lists"].get(l, []) for l in range(n_lists)])
    return o


def query_arrays(case):
    import numpy as np

    q = case["query"]
    return (np.array([0, len(q)], np.int32), np.array([t for t, _ in q], np.uint16),
            np.array([w for _, w in q], np.float32))
