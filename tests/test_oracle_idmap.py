"""Pins the oracle's IDMapIndex (id_map_index.cpp:29-75, id_map_index.h:27-75)
with hand-derived known answers. The index: docs d0 {0: 2}, d1 {1: 3},
d2 {0: 1} added with external ids 100, 7, 55; posting list 0 holds one
cluster [d0, d2] (summary {0: 2}), list 1 one cluster [d1]. The query is
{0: 1} with cut 1, so only list 0 is visited: d0 scores 2, d2 scores 1, and
d1 (ext 7) is never visited by the traversal.

Each case is chosen so that a plausible mistake changes the labels:
  * selectors evaluated on internal ids instead of external ones;
  * the exact-match gate counting only the ids the map knows (the reference's
    wrapper keeps the delegate's size(), unknown ids included);
  * labels left internal."""
import numpy as np
import pytest

from _hand_cases import build_oracle

I2E = [100, 7, 55]
CASE = dict(dim=8, docs=[{0: 2.0}, {1: 3.0}, {0: 1.0}],
            lists={0: [([0, 2], {0: 2.0})], 1: [([1], {1: 3.0})]})
Q = (np.array([0, 1], np.int32), np.array([0], np.uint16), np.array([1.0], np.float32))


def _bitmap(ids, nbits):
    words = np.zeros((nbits + 31) // 32, np.uint32)
    for i in ids:
        words[i >> 5] |= np.uint32(1 << (i & 31))
    return words


# (selector, negate, k) -> (external labels, distances)
KNOWN = [
    # no selector: d0 (2), d2 (1), mapped to 100, 55
    (None, False, 2, [100, 55], [2.0, 1.0]),
    # {7, 999}: size 2 > k = 1, so the traversal runs with the filter; only
    # ext 7 (d1) passes and it is never visited -> padding. Gating on the
    # mapped size (1) would take the exact path and return 7 (score 0).
    (("set", [7, 999]), False, 1, [-1], [-1.0]),
    # the same selector at k = 2: size 2 <= 2 -> exact path over the mapped
    # ids [1]: d1 scores 0 against the query -> [7, pad]
    (("set", [7, 999]), False, 2, [7, -1], [0.0, -1.0]),
    # ArrayIDSelector {55}: size 1 <= 2 -> exact path over [2]: d2 scores 1
    (("array", [55]), False, 2, [55, -1], [1.0, -1.0]),
    # {100, 55, 7}: size 3 > 2 -> traversal, both visited docs pass on their
    # external ids. Evaluated on internal ids (0, 2) none would pass.
    (("set", [100, 55, 7]), False, 2, [100, 55], [2.0, 1.0]),
    # NotIDSelector(Set{100}): not enumerable; d0 (ext 100) is excluded
    (("set", [100]), True, 2, [55, -1], [1.0, -1.0]),
    # a bitmap over external ids with bit 55 set: only d2 passes
    (("bitmap", _bitmap([55], 128), 128), False, 2, [55, -1], [1.0, -1.0]),
]


@pytest.mark.parametrize("tie_mode", [0, 1])
@pytest.mark.parametrize("i", range(len(KNOWN)))
def test_idmap_known_answers(i, tie_mode):
    sel, neg, k, want_l, want_d = KNOWN[i]
    o = build_oracle(CASE, tie_mode)
    d, lab = o.search_idmap(I2E, *Q, k, cut=1, hf=1.0, selector=sel, negate=neg)
    assert list(lab[0]) == want_l
    assert list(d[0]) == want_d


def test_idmap_without_selector_is_the_delegate_relabelled():
    """On a built corpus: IDMap labels are the delegate's mapped through i2e,
    distances identical (no selector)."""
    from _oracle import OracleIndex
    from test_oracle import corpus

    ip, ix, v = corpus(1500, 512, 20, seed=31)
    o = OracleIndex(512, 1, 150, 15, 0.4, 2, 0.0, float(v.max()), tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    i2e = (np.random.default_rng(3).permutation(1500) * 3 + 11).astype(np.int32)
    qip, qix, qv = corpus(40, 512, 10, seed=32)
    d0, l0 = o.search(qip, qix, qv, 10, cut=5, hf=0.9)
    d1, l1 = o.search_idmap(i2e, qip, qix, qv, 10, cut=5, hf=0.9)
    assert (d0 == d1).all()
    assert (l1 == np.where(l0 >= 0, i2e[np.maximum(l0, 0)], -1)).all()
