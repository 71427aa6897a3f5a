"""The oracle against tests/golden/ref_fixtures.json: digests and values the
reference's own build-side code produced (tests/golden/make_ref_fixtures.py),
so the oracle stays pinned where the reference tree is absent.
  * write_index bytes of SEIS / SESQ indexes (libstdc++ tie mode), and of the id
    tie mode for FP32, whose corpora here have no value ties that reach the
    layout -- the oracle's bytes hash to the reference's;
  * InvertedIndex lists + maxima, byte for byte;
  * 8-bit and 16-bit SQ codes and decoded scores;
  * summary scores (score_summaries_transposed) of seeded query / list pairs."""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import _oracle as O
from _oracle import OracleIndex, OracleInverted

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from golden_cases import (BUILD_CASES, INVERTED_CASES, SQ_RANGES, build_corpus, inverted_corpus,  # noqa: E402
                          sq_values, summary_digest, summary_pairs)

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_fixtures.json")) as _f:
    FIX = json.load(_f)


def _digest(b):
    return hashlib.sha256(b).hexdigest(), len(b)


def _case_id(c):
    return f"sq{c['sq']}-n{c['n']}-d{c['dim']}-l{c['lam']}-b{c['beta']}"


def test_fixture_covers_every_case():
    assert len(FIX["build"]) == len(BUILD_CASES) and len(FIX["inverted"]) == len(INVERTED_CASES)
    for fx, case in zip(FIX["build"], BUILD_CASES):
        assert {k: fx[k] for k in case} == case


@pytest.mark.parametrize("idx", range(len(BUILD_CASES)), ids=[_case_id(c) for c in BUILD_CASES])
def test_build_bytes_hash_to_reference(idx):
    case, fx = BUILD_CASES[idx], FIX["build"][idx]
    ip, ix, v = build_corpus(case)
    vmin, vmax = float(v.min()), float(v.max())
    o = OracleIndex(case["dim"], case["sq"], case["lam"], case["beta"], case["alpha"], case["seed"], vmin, vmax,
                    tie_mode=0)
    o.add(ip, ix, v)
    o.build()
    assert _digest(o.serialize()) == (fx["sha256"], fx["size"])


def test_id_tie_mode_hashes_to_reference_without_ties():
    case, fx = BUILD_CASES[0], FIX["build"][0]  # FP32: no value ties in the layout
    assert case["sq"] == 0
    ip, ix, v = build_corpus(case)
    o = OracleIndex(case["dim"], 0, case["lam"], case["beta"], case["alpha"], case["seed"], tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    assert _digest(o.serialize()) == (fx["sha256"], fx["size"])


@pytest.mark.parametrize("idx", range(len(INVERTED_CASES)))
def test_inverted_lists_hash_to_reference(idx):
    case, fx = INVERTED_CASES[idx], FIX["inverted"][idx]
    ip, ix, v = inverted_corpus(case)
    o = OracleInverted(case["dim"])
    o.add(ip, ix, v)
    o.build()
    assert _digest(o.serialize()) == (fx["sha256"], fx["size"])


def test_sq_codes_and_decode_match_reference():
    vals = sq_values()
    for (lo, hi), fx in zip(SQ_RANGES, FIX["sq_encode"]):
        codes = np.empty(len(vals), np.uint8)
        O.lib.oracle_sq_encode(ctypes.c_float(lo), ctypes.c_float(hi), vals.ctypes.data_as(O.P),
                               ctypes.c_size_t(len(vals)), codes.ctypes.data_as(O.P))
        assert codes[:64].tolist() == fx["first_codes"]
        assert hashlib.sha256(codes.tobytes()).hexdigest() == fx["codes_sha256"]
    for fx in FIX["sq_decode"]:
        d = O.lib.oracle_sq_decode(*fx["ingest"], *fx["query"], fx["score"])
        assert int(np.float32(d).view(np.uint32)) == fx["decoded_bits"]


def test_sq16_codes_and_decode_match_reference():
    vals = sq_values()
    for (lo, hi), fx in zip(SQ_RANGES, FIX["sq_encode16"]):
        codes = np.empty(len(vals), np.uint16)
        O.lib.oracle_sq_encode16(ctypes.c_float(lo), ctypes.c_float(hi), vals.ctypes.data_as(O.P),
                                 ctypes.c_size_t(len(vals)), codes.ctypes.data_as(O.P))
        assert codes[:64].tolist() == fx["first_codes"]
        assert hashlib.sha256(codes.tobytes()).hexdigest() == fx["codes_sha256"]
    for fx in FIX["sq_decode16"]:
        d = O.lib.oracle_sq_decode16(*fx["ingest"], *fx["query"], fx["score"])
        assert int(np.float32(d).view(np.uint32)) == fx["decoded_bits"]


@pytest.mark.parametrize("idx", range(len(BUILD_CASES)), ids=[_case_id(c) for c in BUILD_CASES])
def test_summary_scores_hash_to_reference(idx):
    """The oracle's summary scores (libstdc++ tie mode build, which is the
    reference's layout) for every (query, list) pair hash to the reference's."""
    case, fx = BUILD_CASES[idx], FIX["summaries"][idx]
    ip, ix, v = build_corpus(case)
    vmin, vmax = float(v.min()), float(v.max())
    o = OracleIndex(case["dim"], case["sq"], case["lam"], case["beta"], case["alpha"], case["seed"], vmin, vmax,
                    tie_mode=0)
    o.add(ip, ix, v)
    o.build()
    qip, qix, qv, lists = summary_pairs(case)
    assert len(lists) == fx["pairs"]
    sc = o.score_summaries(qip, qix, qv, lists)
    assert sc[0].view(np.uint32).tolist() == fx["first_bits"]
    assert summary_digest(sc) == fx["sha256"]
