"""Writes tests/golden/ref_fixtures.json: digests and values produced by the
reference's own build-side code, compiled from its sources into
oracle/_ref/libnsparse_ref.so by oracle/Makefile. Run it where that library is
built:

    make -C oracle && python tests/golden/make_ref_fixtures.py

The fixture holds data only:
  * the SHA-256 and length of write_index bytes of SEIS / SESQ indexes built by
    the reference from seeded generator corpora (the inputs are regenerated
    from the seeds by the test), for the cases of test_oracle.py;
  * the same for the posting lists + per-term maxima of InvertedIndex::build;
  * 8-bit and 16-bit SQ codes of fixed values under three ranges, and decoded
    scores;
  * per build case, the summary scores InvertedListClusters::
    score_summaries_transposed gives seeded queries against their lists
    (golden_cases.summary_pairs): SHA-256 of every pair's cluster count
    (int32) and scores (float32 bits), and the first pair's scores.
tests/test_golden.py checks the oracle against it without the reference.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import _oracle as O  # noqa: E402
from golden_cases import (BUILD_CASES, INVERTED_CASES, SQ_RANGES, build_corpus, inverted_corpus,  # noqa: E402
                          sq_values, summary_digest, summary_pairs)


def digest(b):
    return {"sha256": hashlib.sha256(b).hexdigest(), "size": len(b)}


def main():
    if O.ref is None:
        raise SystemExit("oracle/_ref/libnsparse_ref.so is not built here")
    out = {"build": [], "inverted": [], "sq_encode": [], "sq_decode": [], "summaries": [], "sq_encode16": [],
           "sq_decode16": []}
    for case in BUILD_CASES:
        ip, ix, v = build_corpus(case)
        vmin, vmax = float(v.min()), float(v.max())
        b = O.ref_build_bytes(case["dim"], case["sq"], ip, ix, v, case["lam_resolved"], case["beta_resolved"],
                              case["alpha"], case["seed"], vmin, vmax)
        out["build"].append(dict(case, **digest(b)))
    for case in BUILD_CASES:
        ip, ix, v = build_corpus(case)
        vmin, vmax = float(v.min()), float(v.max())
        qip, qix, qv, lists = summary_pairs(case)
        sc = O.ref_score_summaries(case["dim"], case["sq"], (ip, ix, v), case["lam_resolved"],
                                   case["beta_resolved"], case["alpha"], case["seed"], (qip, qix, qv), lists, vmin,
                                   vmax)
        out["summaries"].append({"pairs": len(lists), "sha256": summary_digest(sc),
                                 "first_bits": sc[0].view(np.uint32).tolist()})
    for case in INVERTED_CASES:
        ip, ix, v = inverted_corpus(case)
        out["inverted"].append(dict(case, **digest(O.ref_inverted_bytes(case["dim"], ip, ix, v))))
    vals = sq_values()
    for lo, hi in SQ_RANGES:
        codes = np.empty(len(vals), np.uint8)
        O.ref.ref_sq_encode(ctypes.c_float(lo), ctypes.c_float(hi), vals.ctypes.data_as(O.P),
                            ctypes.c_size_t(len(vals)), codes.ctypes.data_as(O.P))
        out["sq_encode"].append({"vmin": lo, "vmax": hi, "codes_sha256": hashlib.sha256(codes.tobytes()).hexdigest(),
                                 "first_codes": codes[:64].tolist()})
    for s in [0.0, 1.0, 12345.0, 3.5e6]:
        d = O.ref.ref_sq_decode(0.01, 43.9, 0.02, 20.0, s)
        out["sq_decode"].append({"ingest": [0.01, 43.9], "query": [0.02, 20.0], "score": s,
                                 "decoded_bits": int(np.float32(d).view(np.uint32))})
    for lo, hi in SQ_RANGES:
        codes = np.empty(len(vals), np.uint16)
        O.ref.ref_sq_encode16(ctypes.c_float(lo), ctypes.c_float(hi), vals.ctypes.data_as(O.P),
                              ctypes.c_size_t(len(vals)), codes.ctypes.data_as(O.P))
        out["sq_encode16"].append({"vmin": lo, "vmax": hi,
                                   "codes_sha256": hashlib.sha256(codes.tobytes()).hexdigest(),
                                   "first_codes": codes[:64].tolist()})
    for s in [0.0, 1.0, 12345.0, 3.5e6, 4.0e12]:
        d = O.ref.ref_sq_decode16(0.01, 43.9, 0.02, 20.0, s)
        out["sq_decode16"].append({"ingest": [0.01, 43.9], "query": [0.02, 20.0], "score": s,
                                   "decoded_bits": int(np.float32(d).view(np.uint32))})
    with open(os.path.join(HERE, "ref_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(out['build'])} build, {len(out['inverted'])} inverted digests")


if __name__ == "__main__":
    main()
