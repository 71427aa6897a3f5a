#!/usr/bin/env python3
"""Batched SEISMIC-SQ search on one MI355X: QPS, roofline, recall, plain path
and the host-CPU baseline, printed as one JSON line.

Workload (defaults): 8,841,823 synthetic docs (vocab 30522, Zipf 1.1 terms
without repeats, doc nnz lognormal median 120 clipped 8..600, FP32
lognormal(0, 0.7) weights), uint8 scalar quantization over the corpus min/max,
lambda=6000, beta=600, alpha=0.4; 10k queries with nnz log-uniform over 5..300
and k in {10, 100} at random, cut=10, heap_factor=0.8. One step = one
batched search launch over all 10k queries, queries and index resident in HBM.

Beside it: the plain drop-in path (one query per launch, synchronous), recall@10
against exact FP32 brute force on a query sample (computed on the GPU), and the
CPU restatement of the reference search timed on this host over a bounded query
sample on the very same index.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-sparse-cpp_amd"))

METRIC = "QPS at recall@10≥0.95, 8.8M-doc SEISMIC-SQ: batched vs plain; p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    """Host CPU model name and socket count from /proc/cpuinfo."""
    name, sockets = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key = key.strip()
                if key == "model name" and name is None:
                    name = val.strip()
                elif key == "physical id":
                    sockets.add(val.strip())
    except OSError:
        pass
    return name, max(len(sockets), 1)


def algorithmic_bytes(stats, probe_bytes=None):
    """Bytes one query's traversal reads, per row of traversal counters in the
    order lists, terms_found, csc_entries, clusters_tested, clusters_opened,
    docs_read, docs_scored, nnz_scored[, probe_bytes], uint8 values: per
    visited list the binary-search probes over its term ids (q_len * 2 B *
    ceil(log2 T)); per query entry found in the summaries its term_ptr pair
    (8 B) and entries (2 B cluster + 1 B value); per prune test the cluster's
    offsets pair (8 B); per doc id read 4 B; per row scored its indptr pair
    (8 B) and nnz * (2 B term + 1 B code). With the oracle's counters (the
    reference's traversal, whose visited set skips a doc already scored in an
    earlier list) these are the reference's bytes; with the device's, they
    include the rows it scores again for such docs."""
    s = stats.astype(np.float64)
    probe = s[:, 8] if probe_bytes is None else np.asarray(probe_bytes, np.float64)
    return (probe + 8 * s[:, 1] + 3 * s[:, 2] + 8 * s[:, 3] + 4 * s[:, 5] + 8 * s[:, 6] + 3 * s[:, 7])


def serving_bench(args, nsparse, index, n_docs, threads):
    """Continuous batching on the resident index: Poisson arrivals at each rate,
    k in {10, 100}, 10% of requests carrying a selector bitmap that excludes a
    random 20% of ids; p50 / p99 latency (arrival to result on the host) and
    sustained QPS (requests / (last completion - first arrival))."""
    from nsparse import synth

    rates = [float(r) for r in args.serve_rates.split(",") if r and r != "none"]
    if not rates:
        return None
    rng = np.random.default_rng(3)
    srv = nsparse.Server(index, max_batch=args.serve_batch, streams=args.serve_streams,
                         scratch_budget=args.serve_budget_mb << 20, k_max=100, cut=args.cut, heap_factor=args.hf)
    for _ in range(8):
        keep = rng.random(n_docs) >= 0.2
        b = np.packbits(keep.astype(np.uint8), bitorder="little")
        srv.add_bitmap(np.pad(b, (0, (-len(b)) % 4)).view(np.uint32))
    out = []
    for rate in rates:
        n = int(min(max(rate * args.serve_seconds, 2000), 60000))
        qip, qix, qv = synth.make_corpus(synth.query_spec(seed=100 + int(rate), log_uniform=True), n,
                                         threads=threads)
        r2 = np.random.default_rng(int(rate))
        ks = np.where(r2.random(n) < 0.5, 10, 100).astype(np.int32)
        bids = np.where(r2.random(n) < 0.1, r2.integers(0, 8, n), -1).astype(np.int32)
        arrivals = np.cumsum(r2.exponential(1.0 / rate, n))
        before = srv.stats()
        _, _, lat = srv.replay(arrivals, qip, qix, qv, ks, bids)
        after = srv.stats()
        done = arrivals + lat * 1e-6
        lat_s = np.sort(lat) / 1e3
        pct = lambda f: float(lat_s[min(n - 1, int(math.floor(f * n)))])  # noqa: E731
        batches = after["batches"] - before["batches"]
        row = {"offered_qps": rate, "requests": n, "sustained_qps": n / float(done.max() - arrivals[0]),
               "p50_ms": pct(0.5), "p99_ms": pct(0.99), "batches": batches, "mean_batch": n / max(batches, 1),
               "max_concurrent_batches": after["max_concurrent_batches"],
               "deferred_by_budget": after["deferred_by_budget"] - before["deferred_by_budget"]}
        log(f"serve {rate:,.0f}/s: sustained {row['sustained_qps']:,.0f} QPS, p50 {row['p50_ms']:.2f} ms, "
            f"p99 {row['p99_ms']:.2f} ms, {batches} batches (mean {row['mean_batch']:.1f})")
        out.append(row)
    srv.close()
    return {"config": {"max_batch": args.serve_batch, "streams": args.serve_streams,
                       "scratch_budget_mb": args.serve_budget_mb, "k": "10 or 100", "selectors": "10%, 20% excluded",
                       "arrivals": "Poisson"}, "rates": out}


def measured_traffic(workload):
    """HBM bytes per launch of this workload as measured by the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarized under profiles/ (scaled by a copy
    probe there); None when no summary covers the workload."""
    import glob
    import re

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"traffic_bytes_per_launch\[(.+)\] = (\d+)", line.strip())
                if m and m.group(1) == workload:
                    return float(m.group(2)), os.path.relpath(path, ROOT)
    return None, None


def build_bytes(n_docs, nnz, st, v=1):
    """Bytes the reference's build reads and writes, uint8 values (v = 1): the
    forward index read once for list construction ((N + 1) * 4 + nnz * (2 +
    v)), the postings written and read back (2 * nnz * (4 + v)), every kept
    posting's doc row gathered twice, for the assignment and the summaries
    (2 * (8 * kept + row_nnz * (2 + v))), and one centroid-index entry (2 B
    cluster id + v) per (doc term, centroid posting) pair the assignment
    scores (kmeans_utils.cpp:107-156). Returns (total, the pairs' share)."""
    fwd = (n_docs + 1) * 4 + nnz * (2 + v)
    postings = 2 * nnz * (4 + v)
    rows = 2 * (8 * st["kept_postings"] + st["kept_row_nnz"] * (2 + v))
    pairs = st["assign_pairs"] * (2 + v)
    return float(fwd + postings + rows + pairs), float(pairs)


def read_bandwidth(nsparse, dev, nbytes=4 << 30, reps=5):
    """Rate of a streaming read of device memory (GB/s): the library's read
    probe (non-temporal 16-byte loads, XOR of the words; the fastest read shape
    measured on this device) over a buffer far larger than the last-level
    cache, timed with HIP events on the stream it runs on."""
    import ctypes

    import torch

    buf = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    out = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)

    def go():
        rc = nsparse._lib.lib.nsparse_testing_read_probe(ctypes.c_void_p(buf.data_ptr()), ctypes.c_size_t(nbytes),
                                                        ctypes.c_void_p(out.data_ptr()), sp)
        if rc != 0:
            raise RuntimeError(nsparse._lib.lib.nsparse_last_error().decode())

    go()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        go()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del buf, out
    torch.cuda.empty_cache()
    return nbytes / (ms * 1e-3) / 1e9


def copy_bandwidth(dev, nbytes=2 << 30, reps=5):
    """Rate of a plain device copy (bytes read + bytes written per second, GB/s)
    over buffers far larger than the last-level cache: the measured ceiling
    beside the published HBM peak."""
    import torch

    src = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def inverted_bench(args, nsparse, docs, queries, truth, threads):
    """The exact InvertedIndex on the same FP32 corpus: GPU build, then the
    drop-in search (host queries, as the reference API takes them, k = 10) of a
    query sample; recall@10 against the FP32 ground truth; algorithmic bytes =
    every posting of every query term (doc id + value, 8 B) plus one n_docs
    float score row written and read per query. Beside it, the oracle's
    restatement of the reference's MaxScore traversal on all host threads."""
    n = min(args.inverted_queries, args.queries)
    if n <= 0:
        return None
    import torch

    inv = nsparse.index_factory(30522, "inverted")
    t0 = time.time()
    inv.add(args.docs, *docs)
    inv.build()
    torch.cuda.synchronize()
    build_s = time.time() - t0
    qip = np.ascontiguousarray(queries[0][: n + 1])
    qix, qv = queries[1][: qip[-1]], queries[2][: qip[-1]]
    inv.search(min(16, n), qip[: min(16, n) + 1], qix, qv, 10)
    t0 = time.perf_counter()
    d, lab = inv.search(n, qip, qix, qv, 10)
    wall = time.perf_counter() - t0
    df = np.bincount(docs[1], minlength=30522).astype(np.int64)
    postings = int(df[qix].sum())
    alg = postings * 8 + n * args.docs * 8
    n_rec = min(n, len(truth))
    hits = [len(set(int(x) for x in lab[i] if x >= 0) & truth[i]) / len(truth[i]) for i in range(n_rec) if truth[i]]
    out = {"queries": n, "k": 10, "qps": n / wall, "build_s": build_s, "postings_per_query": postings / n,
           "algorithmic_bytes_per_query": alg / n, "achieved_gbps": alg / wall / 1e9,
           "frac": alg / wall / 1e9 / HBM_PEAK_GBS, "recall_at_10": float(np.mean(hits)) if hits else None,
           "recall_queries": n_rec}
    log(f"inverted (exact): GPU build {build_s:.1f}s; {out['qps']:,.0f} QPS over {n} queries (drop-in, k=10), "
        f"{postings / n / 1e6:.1f}M postings/query, {out['achieved_gbps']:,.0f} GB/s algorithmic, "
        f"recall@10 {out['recall_at_10']}")
    del inv
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle

        t0 = time.time()
        ref = _oracle.OracleInverted(30522, tie_mode=0)
        ref.add(*docs)
        ref.build()
        ref_build_s = time.time() - t0
        done, t_cpu, same = 0, 0.0, 0
        chunk = 16 * threads
        while t_cpu < args.cpu_seconds / 2 and done < n:
            hi = min(done + chunk, n)
            sip = (qip[done:hi + 1] - qip[done]).astype(np.int32)
            t1 = time.perf_counter()
            cd, cl = ref.search(sip, qix[qip[done]:qip[hi]], qv[qip[done]:qip[hi]], 10, threads=threads)
            t_cpu += time.perf_counter() - t1
            for r in range(hi - done):
                same += int(set(cl[r].tolist()) == set(lab[done + r].tolist()))
            done = hi
        out["cpu_baseline"] = {"value": done / t_cpu, "unit": "queries/s", "cores": threads, "kind": "port",
                               "sample": f"first {done} of the {n} inverted-benchmark queries, k=10, {chunk} per "
                                         "call, OpenMP over queries",
                               "oracle_build_s": ref_build_s, "label_sets_equal_frac": same / max(done, 1)}
        log(f"inverted cpu (MaxScore, {threads} threads): {done / t_cpu:,.1f} QPS over {done} queries; "
            f"label sets equal to the GPU's {same}/{done}; oracle build {ref_build_s:.1f}s")
        del ref
    return out


def same_topk_up_to_ties(d_ref, l_ref, d_got, l_got):
    """Top-k identical apart from exact score ties: the same scores in the same
    order, and the same doc ids at every score above the k-th (docs tied with
    the k-th may be any of the tied ones, and equal scores may list their docs
    in any order)."""
    if not np.array_equal(d_ref, d_got):
        return False
    kth = d_ref[-1]
    above = d_ref > kth
    if not np.array_equal(np.sort(l_ref[above]), np.sort(l_got[above])):
        return False
    return bool(np.array_equal(d_ref == -1.0, d_got == -1.0))


def exact_truth(args, nsparse, index, docs, queries, qix, qv, vmin, vmax, n_rec, dev, stream):
    """Exact top-10 (score desc, id asc, positive scores) of the first n_rec
    queries on the GPU: FP32 on the unquantized corpus, and on the index's own
    uint8 codes with the query quantized as the index does it."""
    import ctypes

    import torch

    d_ip = torch.from_numpy(docs[0]).to(dev)
    d_ix = torch.from_numpy(docs[1].view(np.int16)).to(dev)
    d_v = torch.from_numpy(docs[2]).to(dev)
    scores = torch.empty(args.docs, dtype=torch.float32, device=dev)
    truth = []
    for i in range(n_rec):
        a, b = int(queries[0][i]), int(queries[0][i + 1])
        nsparse._lib.check(nsparse._lib.lib.nsparse_exact_scores_device(
            args.docs, d_ip.data_ptr(), d_ix.data_ptr(), d_v.data_ptr(), 30522, b - a,
            ctypes.c_void_p(qix.data_ptr() + 2 * a), ctypes.c_void_p(qv.data_ptr() + 4 * a), scores.data_ptr(),
            ctypes.c_void_p(stream.cuda_stream)))
        top_v, top_i = torch.topk(scores, 64)
        tv, ti = top_v.cpu().numpy(), top_i.cpu().numpy().astype(np.int64)
        keep = tv > 0
        tv, ti = tv[keep], ti[keep]
        order = np.lexsort((ti, -tv))[:10]
        truth.append(set(ti[order].tolist()))
    # the same brute force on the index's own uint8 codes (query quantized like the
    # index does it): the quantization ceiling, separate from SEISMIC's pruning
    _, _, codes = index.get_vectors()
    d_v = torch.from_numpy(codes).to(dev).to(torch.float32)
    del codes
    scale = np.float32(255.0) / (np.float32(vmax) - np.float32(vmin))
    qx = (queries[2].astype(np.float32) - np.float32(vmin)) * scale
    qx = np.minimum(np.maximum(qx, np.float32(0.0)), np.float32(255.0))
    qcodes = torch.from_numpy(np.floor(qx.astype(np.float64) + 0.5).astype(np.float32)).to(dev)
    truth_sq = []
    for i in range(n_rec):
        a, b = int(queries[0][i]), int(queries[0][i + 1])
        nsparse._lib.check(nsparse._lib.lib.nsparse_exact_scores_device(
            args.docs, d_ip.data_ptr(), d_ix.data_ptr(), d_v.data_ptr(), 30522, b - a,
            ctypes.c_void_p(qix.data_ptr() + 2 * a), ctypes.c_void_p(qcodes.data_ptr() + 4 * a), scores.data_ptr(),
            ctypes.c_void_p(stream.cuda_stream)))
        top_v, top_i = torch.topk(scores, 64)
        tv, ti = top_v.cpu().numpy(), top_i.cpu().numpy().astype(np.int64)
        keep = tv > 0
        tv, ti = tv[keep], ti[keep]
        order = np.lexsort((ti, -tv))[:10]
        truth_sq.append(set(ti[order].tolist()))
    del d_ip, d_ix, d_v, scores, qcodes

    return truth, truth_sq


def fp32_truth(nsparse, docs, queries, n_rec, dev, stream):
    """Exact FP32 top-10 (score desc, id asc, positive scores) of the first
    n_rec queries on the unquantized corpus, scored on the GPU."""
    import ctypes

    import torch

    n_docs = len(docs[0]) - 1
    d_ip = torch.from_numpy(docs[0]).to(dev)
    d_ix = torch.from_numpy(docs[1].view(np.int16)).to(dev)
    d_v = torch.from_numpy(docs[2]).to(dev)
    q_ix = torch.from_numpy(queries[1].view(np.int16)).to(dev)
    q_v = torch.from_numpy(queries[2]).to(dev)
    scores = torch.empty(n_docs, dtype=torch.float32, device=dev)
    truth = []
    for i in range(n_rec):
        a, b = int(queries[0][i]), int(queries[0][i + 1])
        nsparse._lib.check(nsparse._lib.lib.nsparse_exact_scores_device(
            n_docs, d_ip.data_ptr(), d_ix.data_ptr(), d_v.data_ptr(), 30522, b - a,
            ctypes.c_void_p(q_ix.data_ptr() + 2 * a), ctypes.c_void_p(q_v.data_ptr() + 4 * a), scores.data_ptr(),
            ctypes.c_void_p(stream.cuda_stream)))
        top_v, top_i = torch.topk(scores, 64)
        tv, ti = top_v.cpu().numpy(), top_i.cpu().numpy().astype(np.int64)
        keep = tv > 0
        tv, ti = tv[keep], ti[keep]
        truth.append(set(ti[np.lexsort((ti, -tv))[:10]].tolist()))
    return truth


def c1_reference_order(args, nsparse, threads, dev, stream):
    """The 1M-doc SEISMIC-SQ configuration (lambda 6000, beta 600, alpha 0.4;
    queries of nnz lognormal median 40, k = 10, cut and heap_factor as the
    benchmark's): the index built in the reference's own tie order by the
    oracle (its bytes are the reference build's) on all host threads -- the
    CPU build baseline -- beside the GPU build of the same corpus. Then
    recall@10 against exact FP32 brute force for the reference path (the
    oracle's traversal with libstdc++ ties) on the reference-order index, for
    the device on that same index (read from its bytes), and for the device on
    its own build (ties broken by id)."""
    if args.c1_docs <= 0 or args.no_cpu:
        return None
    import torch

    from nsparse import synth

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle

    n, nq = args.c1_docs, args.c1_queries
    docs = synth.make_corpus(synth.doc_spec(seed=0), n, threads=threads)
    queries = synth.make_corpus(synth.query_spec(seed=1), nq, threads=threads)
    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    desc = (f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda={args.lam}|beta={args.beta}|"
            f"alpha={args.alpha}|seed=0")
    gpu = nsparse.index_factory(30522, desc)
    gpu.add(n, *docs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gpu.build()
    torch.cuda.synchronize()
    gpu_build_s = time.perf_counter() - t0
    o = _oracle.OracleIndex(30522, 1, args.lam, args.beta, args.alpha, 0, np.float32(vmin), np.float32(vmax),
                            tie_mode=0)
    o.add(*docs)
    t0 = time.perf_counter()
    o.build(threads=threads)
    cpu_build_s = time.perf_counter() - t0
    log(f"C1 {n} docs: GPU build {gpu_build_s:.2f}s, oracle build in the reference's tie order "
        f"{cpu_build_s:.1f}s on {threads} threads ({cpu_build_s / gpu_build_s:.0f}x)")
    ref_dev = o.to_library_index(nsparse)
    qip, qix, qv = queries
    p = nsparse.SeismicSearchParameters(args.cut, args.hf)
    d_rd, l_rd = ref_dev.search(nq, qip, qix, qv, 10, p)
    d_g, l_g = gpu.search(nq, qip, qix, qv, 10, p)
    d_c, l_c = o.search(qip, qix, qv, 10, cut=args.cut, hf=args.hf, threads=threads)
    o1 = _oracle.OracleIndex(30522, 1, args.lam, args.beta, args.alpha, 0, np.float32(vmin), np.float32(vmax),
                             tie_mode=1)
    o1.copy_from(o)
    d_1, l_1 = o1.search(qip, qix, qv, 10, cut=args.cut, hf=args.hf, threads=threads)
    exact = float(np.mean((l_rd == l_1).all(axis=1) & (d_rd == d_1).all(axis=1)))
    up_to_ties = float(np.mean([same_topk_up_to_ties(d_c[i], l_c[i], d_rd[i], l_rd[i]) for i in range(nq)]))
    n_rec = min(args.c1_recall_queries, nq)
    truth = fp32_truth(nsparse, docs, queries, n_rec, dev, stream)

    def recall(lab):
        hits = [len(set(int(x) for x in lab[i] if x >= 0) & t) / len(t) for i, t in enumerate(truth) if t]
        return float(np.mean(hits))

    r_ref, r_rd, r_g = recall(l_c), recall(l_rd), recall(l_g)
    log(f"C1 recall@10 over {n_rec} queries: reference path {r_ref:.4f}, GPU on the reference-order index "
        f"{r_rd:.4f} ({r_rd - r_ref:+.4f}), GPU on its own build {r_g:.4f} ({r_g - r_ref:+.4f}); GPU rows equal "
        f"to the oracle (id ties) {exact:.4f}, to the reference path up to ties {up_to_ties:.4f}")
    del gpu, ref_dev, o, o1
    return {"docs": n, "queries": nq, "k": 10, "cut": args.cut, "heap_factor": args.hf,
            "gpu_build_s": gpu_build_s,
            "cpu_build_baseline": {"value": cpu_build_s, "unit": "s", "cores": threads, "kind": "port",
                                   "sample": f"the whole {n}-doc build, the oracle in the reference's tie order "
                                             "(its write_index bytes are the reference build's), OpenMP over "
                                             "posting lists"},
            "recall_queries": n_rec, "recall_reference_path": r_ref,
            "recall_gpu_on_reference_order_index": r_rd, "recall_gpu_on_own_build": r_g,
            "gpu_minus_reference_same_index": r_rd - r_ref, "gpu_own_build_minus_reference": r_g - r_ref,
            "gpu_rows_equal_oracle_id_ties": exact, "gpu_rows_equal_reference_up_to_ties": up_to_ties}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--docs", type=int, default=8_841_823)
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--lam", type=int, default=6000)
    ap.add_argument("--beta", type=int, default=600)
    ap.add_argument("--alpha", type=float, default=0.4)
    ap.add_argument("--cut", type=int, default=10)
    ap.add_argument("--hf", type=float, default=0.8)
    ap.add_argument("--plain-queries", type=int, default=300)
    ap.add_argument("--recall-queries", type=int, default=1000,
                    help="queries whose recall@10 is measured against exact FP32 brute force")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--sweep", default="", help="comma-separated cut:heap_factor pairs to time and score")
    ap.add_argument("--serve-rates", default="1000,10000,50000,200000",
                    help="serving loop: offered QPS levels (the workload's 1k, 10k and 50k/s, and a load near the "
                         "loop's capacity)")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-batch", type=int, default=1024)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--serve-budget-mb", type=int, default=256)
    ap.add_argument("--rescore-candidates", type=int, default=20,
                    help="two-stage operating point: SQ candidates per query re-scored in FP32 (0 skips)")
    ap.add_argument("--rescore-cut", type=int, default=8, help="two-stage operating point: the SQ search's cut")
    ap.add_argument("--rescore-hf", type=float, default=0.8,
                    help="two-stage operating point: the SQ search's heap_factor")
    ap.add_argument("--inverted-queries", type=int, default=300,
                    help="exact InvertedIndex on the same corpus: drop-in search of this many queries (0 skips)")
    ap.add_argument("--c1-docs", type=int, default=1_000_000,
                    help="reference-order build at this size: CPU build baseline and recall parity (0 skips)")
    ap.add_argument("--c1-queries", type=int, default=10_000)
    ap.add_argument("--c1-recall-queries", type=int, default=1000)
    ap.add_argument("--schedule", default="auto",
                    help='"auto" (autotuner) or a JSON search config, e.g. {"lookup":1,"block":512,"persistent":1}')
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (query-sharded replicas are a "
              "design note only)", file=sys.stderr)
        sys.exit(2)

    import torch

    import nsparse
    from nsparse import synth

    dev = torch.device("cuda:0")
    threads = host_threads()
    t0 = time.time()
    docs = synth.make_corpus(synth.doc_spec(seed=0), args.docs, threads=threads)
    queries = synth.make_corpus(synth.query_spec(seed=1, log_uniform=True), args.queries, threads=threads)
    ks = np.where(np.random.default_rng(2).random(args.queries) < 0.5, 10, 100).astype(np.int32)
    log(f"corpus {args.docs} docs / {int(docs[0][-1])} nnz generated in {time.time() - t0:.1f}s")

    vmin, vmax = float(docs[2].min()), float(docs[2].max())
    desc = (f"seismic_sq,quantizer=8bit|vmin={vmin:.9g}|vmax={vmax:.9g}|lambda={args.lam}|beta={args.beta}|"
            f"alpha={args.alpha}|seed=0")
    index = nsparse.index_factory(30522, desc)
    t0 = time.time()
    index.add(args.docs, *docs)
    add_s = time.time() - t0
    t0 = time.time()
    index.build()
    torch.cuda.synchronize()
    build_s = time.time() - t0
    bst = index.build_stats()
    b_bytes, pair_bytes = build_bytes(args.docs, int(docs[0][-1]), bst)
    build = dict(bst, seconds=build_s, algorithmic_bytes=b_bytes, assign_pair_bytes=pair_bytes,
                 gbps=b_bytes / build_s / 1e9, frac=b_bytes / build_s / 1e9 / HBM_PEAK_GBS,
                 assign_pairs_per_s=bst["assign_pairs"] / build_s)
    log(f"add {add_s:.1f}s, GPU build {build_s:.1f}s: {b_bytes / 1e9:.1f} GB algorithmic "
        f"({build['gbps']:.1f} GB/s), {bst['assign_pairs'] / 1e9:.2f}G assignment pairs "
        f"({build['assign_pairs_per_s'] / 1e9:.1f}G/s), {bst['kept_postings']:,} kept postings, "
        f"{bst['clusters']:,} clusters")

    qip = torch.from_numpy(queries[0]).to(dev)
    qix = torch.from_numpy(queries[1].view(np.int16)).to(dev)
    qv = torch.from_numpy(queries[2]).to(dev)
    d_ks = torch.from_numpy(ks).to(dev)
    k_max = int(ks.max())
    params = nsparse.SeismicSearchParameters(args.cut, args.hf)
    out_d = torch.empty((args.queries, k_max), dtype=torch.float32, device=dev)
    out_l = torch.empty((args.queries, k_max), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    # one step = one replay of the captured batched search; the schedule (lookup
    # table, workgroup width, persistent queue, longest-first order) is the
    # autotuner's pick on this very batch
    t0 = time.time()
    plan = index.search_plan(qip, qix, qv, k_max, ks=d_ks, params=params, out_d=out_d, out_l=out_l,
                             config=None if args.schedule == "auto" else json.loads(args.schedule), stream=stream)
    torch.cuda.synchronize()
    tuning = plan.tuning
    log(f"plan in {time.time() - t0:.1f}s; schedule {plan.config}; tried "
        + ", ".join(f"{c['lookup']}/{c['block']}/{c['persistent']}/{c['longest_first']}={ms:.1f}ms"
                    for c, ms in tuning))
    for _ in range(args.warmup):
        plan.launch(stream)
    torch.cuda.synchronize()
    if index.device_error() != 0:
        raise RuntimeError(f"device search error {index.device_error()}")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        plan.launch(stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    qps = args.queries * args.steps / wall
    log(f"batched: {qps:,.0f} QPS, {wall / args.steps * 1e3:.2f} ms/step, graph {kernel_ms:.2f} ms")
    if args.copy_probe:
        # a plain copy of known size, for scaling rocprofv3 FETCH/WRITE counters
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    labels = out_l.cpu().numpy()
    dists = out_d.cpu().numpy()

    stats = torch.zeros((args.queries, 16), dtype=torch.int64, device=dev)
    index.search_device(qip, qix, qv, k_max, ks=d_ks, params=params, stats=stats, stream=stream)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    dev_bytes = algorithmic_bytes(st)  # per query, by the device's own counters
    bytes_launch = float(dev_bytes.sum())
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    tick_names = ["setup", "summary", "order", "plan", "gather", "score", "replay"]
    tick_means = st[:, 9:16].mean(axis=0) * 10.0 / 1e3  # 100 MHz device wall clock -> us
    log("per-query phase time (us, one workgroup): " + ", ".join(f"{n}={v:.0f}" for n, v in zip(tick_names,
                                                                                               tick_means)))
    log(f"device-counted {bytes_launch / 1e9:.3f} GB/launch -> {achieved:.0f} GB/s; per-query means "
        + ", ".join(f"{n}={v:.0f}" for n, v in zip(["lists", "found", "csc", "tested", "opened", "read", "scored",
                                                      "nnz", "probe_bytes"], st.mean(axis=0))))

    # ---- plain path: one query per launch, synchronous
    n_plain = min(args.plain_queries, args.queries)
    lat = []
    for i in range(n_plain):
        k = int(ks[i])
        od = torch.empty((1, k), dtype=torch.float32, device=dev)
        ol = torch.empty((1, k), dtype=torch.int32, device=dev)
        t1 = time.perf_counter()
        index.search_device(qip[i:i + 2], qix, qv, k, params=params, out_d=od, out_l=ol, stream=stream)
        stream.synchronize()
        lat.append(time.perf_counter() - t1)
        if i < 20:
            assert (ol.cpu().numpy()[0] == labels[i, :k]).all(), "plain and batched results differ"
    lat = np.sort(np.array(lat))
    plain_qps = n_plain / float(lat.sum()) if n_plain else float("nan")
    # percentile as the reference benchmarks define it: sorted[floor(frac * n)]
    pct = lambda f: float(lat[min(len(lat) - 1, int(math.floor(f * len(lat))))] * 1e3) if n_plain else None  # noqa
    if n_plain:
        log(f"plain: {plain_qps:,.0f} QPS, p50 {pct(0.5):.3f} ms, p99 {pct(0.99):.3f} ms")

    # ---- the same queries through Index::search over host arrays (PCIe both ways)
    lat_h = []
    for i in range(n_plain):
        k = int(ks[i])
        a, b = int(queries[0][i]), int(queries[0][i + 1])
        hip_ = np.array([0, b - a], dtype=np.int32)
        t1 = time.perf_counter()
        _, hl = index.search(1, hip_, queries[1][a:b], queries[2][a:b], k, params)
        lat_h.append(time.perf_counter() - t1)
        if i < 20:
            assert (hl[0] == labels[i, :k]).all(), "host-array and batched results differ"
    lat_h = np.sort(np.array(lat_h))
    host_qps = n_plain / float(lat_h.sum()) if n_plain else float("nan")
    pct_h = lambda f: float(lat_h[min(len(lat_h) - 1, int(math.floor(f * len(lat_h))))] * 1e3) if n_plain else None  # noqa
    if n_plain:
        log(f"plain over host arrays: {host_qps:,.0f} QPS, p50 {pct_h(0.5):.3f} ms, p99 {pct_h(0.99):.3f} ms")

    # ---- recall@10 against exact FP32 scores on the unquantized corpus (GPU)
    import ctypes

    n_rec = min(args.recall_queries, args.queries)
    truth, truth_sq = [], []
    if n_rec:
        truth, truth_sq = exact_truth(args, nsparse, index, docs, queries, qix, qv, vmin, vmax, n_rec, dev, stream)

    def recall_of(lab, gt=truth):
        hits = [len(set(int(x) for x in lab[i, :10] if x >= 0) & t) / len(t) for i, t in enumerate(gt) if t]
        return float(np.mean(hits)) if hits else float("nan")

    recall = recall_of(labels)
    recall_sq = recall_of(labels, truth_sq)
    ceiling = float(np.mean([len(a & b) / len(a) for a, b in zip(truth, truth_sq) if a])) if n_rec else float("nan")
    log(f"recall@10 = {recall:.4f} over {n_rec} queries (vs uint8-exact {recall_sq:.4f}; "
        f"uint8-exact vs FP32-exact {ceiling:.4f})")

    # ---- an operating point beyond the reference's arithmetic: the SQ search
    # mines candidates, which are re-scored exactly in FP32 on the device
    rescored = None
    if n_rec and args.rescore_candidates > 0:
        kc = args.rescore_candidates
        d_ip = torch.from_numpy(docs[0]).to(dev)
        d_ix = torch.from_numpy(docs[1].view(np.int16)).to(dev)
        d_v = torch.from_numpy(docs[2]).to(dev)
        cand_d = torch.empty((args.queries, kc), dtype=torch.float32, device=dev)
        cand_l = torch.empty((args.queries, kc), dtype=torch.int32, device=dev)
        fp32 = torch.empty((args.queries, kc), dtype=torch.float32, device=dev)
        err = torch.zeros(1, dtype=torch.int32, device=dev)
        rparams = nsparse.SeismicSearchParameters(args.rescore_cut, args.rescore_hf)
        cplan = index.search_plan(qip, qix, qv, kc, params=rparams, out_d=cand_d, out_l=cand_l, stream=stream)

        def two_stage():
            cplan.launch(stream)
            nsparse._lib.check(nsparse._lib.lib.nsparse_rescore_device(
                0, args.queries, qip.data_ptr(), qix.data_ptr(), qv.data_ptr(), kc, cand_l.data_ptr(), args.docs,
                d_ip.data_ptr(), d_ix.data_ptr(), d_v.data_ptr(), 30522, fp32.data_ptr(), None, None, None,
                err.data_ptr(), ctypes.c_void_p(stream.cuda_stream)))
            return torch.topk(fp32.masked_fill(cand_l < 0, float("-inf")), 10, dim=1)

        two_stage()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            _, top = two_stage()
        e1.record(stream)
        torch.cuda.synchronize()
        r_ms = e0.elapsed_time(e1) / args.steps
        best = torch.gather(cand_l, 1, top).cpu().numpy()
        rescored = {"candidates": kc, "cut": args.rescore_cut, "heap_factor": args.rescore_hf,
                    "qps": args.queries / (r_ms * 1e-3), "ms_per_step": r_ms, "recall_at_10": recall_of(best),
                    "recall_queries": n_rec, "candidate_schedule": cplan.config,
                    "note": f"SQ batched search (cut {args.rescore_cut}, heap_factor {args.rescore_hf}) for {kc} "
                            "candidates per query, their exact FP32 scores from the "
                            "re-scoring kernel, top-10 by FP32 score (torch.topk): an operating point beyond the "
                            "reference's arithmetic"}
        log(f"two-stage: SQ top-{kc} (cut {args.rescore_cut}, hf {args.rescore_hf}) + FP32 re-score: "
            f"{rescored['qps']:,.0f} QPS ({r_ms:.2f} ms/step), "
            f"recall@10 {rescored['recall_at_10']:.4f}")
        del d_ip, d_ix, d_v, cand_d, cand_l, fp32, cplan

    # ---- optional operating-point sweep: batched QPS and recall per (cut, heap_factor)
    sweep = []
    for item in filter(None, args.sweep.split(",")):
        cut, hf = int(item.split(":")[0]), float(item.split(":")[1])
        p = nsparse.SeismicSearchParameters(cut, hf)
        index.search_device(qip, qix, qv, k_max, ks=d_ks, params=p, out_d=out_d, out_l=out_l, stream=stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            index.search_device(qip, qix, qv, k_max, ks=d_ks, params=p, out_d=out_d, out_l=out_l, stream=stream)
        torch.cuda.synchronize()
        sq = args.queries * args.steps / (time.perf_counter() - t1)
        lab = out_l.cpu().numpy()
        rc, rcq = recall_of(lab), recall_of(lab, truth_sq)
        sweep.append({"cut": cut, "heap_factor": hf, "qps": sq, "recall_at_10": rc, "recall_vs_uint8_exact": rcq})
        log(f"sweep cut={cut} hf={hf}: {sq:,.0f} QPS, recall@10 {rc:.4f} (vs uint8-exact {rcq:.4f})")

    # ---- serving loop on the same resident index
    serving = serving_bench(args, nsparse, index, args.docs, threads)

    # ---- the exact InvertedIndex on the same FP32 corpus
    inverted = inverted_bench(args, nsparse, docs, queries, truth, threads)

    # ---- CPU baseline: the reference's search restated in C++ (oracle/),
    # on the same index bytes, all host threads, bounded query sample.
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle

        t1 = time.time()
        ref = _oracle.OracleIndex.load_from_index(nsparse, index, tie_mode=0)
        log(f"oracle copy of the index loaded in {time.time() - t1:.1f}s")
        cpu_labels = np.full((n_rec, 10), -1, np.int64)  # the CPU's top-10 of the recall queries
        ostats = np.zeros((args.queries, 8), np.int64)  # the reference traversal's counters per query
        covered = np.zeros(args.queries, bool)

        def cpu_run(n_threads, seconds, chunk, keep_stats=False):
            """Searches the benchmark queries in order, chunk by chunk, until
            `seconds` of search time; returns (queries, seconds, agree, total)."""
            done, t_cpu, agree, total = 0, 0.0, 0, 0
            while t_cpu < seconds and done < args.queries:
                sl = slice(done, min(done + chunk, args.queries))
                for k in (10, 100):
                    sel = np.nonzero(ks[sl] == k)[0] + sl.start
                    if len(sel) == 0:
                        continue
                    rows = [queries[1][queries[0][j]:queries[0][j + 1]] for j in sel]
                    vals = [queries[2][queries[0][j]:queries[0][j + 1]] for j in sel]
                    sip = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.int32)
                    t1 = time.perf_counter()
                    cd, cl, cs = ref.search(sip, np.concatenate(rows), np.concatenate(vals), k, cut=args.cut,
                                            hf=args.hf, threads=n_threads, with_stats=True)
                    t_cpu += time.perf_counter() - t1
                    if keep_stats:
                        ostats[sel] = cs
                        covered[sel] = True
                    for r, j in enumerate(sel):
                        total += 1
                        agree += int(same_topk_up_to_ties(cd[r], cl[r], dists[j, :k], labels[j, :k]))
                        if j < n_rec:
                            cpu_labels[j] = cl[r, :10]
                done = sl.stop
            return done, t_cpu, agree, total

        # each call carries 128 queries per thread (per k), so OpenMP's
        # dynamic(64) chunks keep every thread busy
        done, t_cpu, agree, total = cpu_run(threads, args.cpu_seconds, 256 * threads, keep_stats=True)
        done1, t_cpu1, _, _ = cpu_run(1, args.cpu_seconds / 3, 8)
        model, sockets = cpu_model()
        simd = "avx512" if _oracle.lib.oracle_set_simd(1) else "generic"
        cpu = {"value": done / t_cpu, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first {done} of the {args.queries} benchmark queries on the same index bytes, "
                         f"{256 * threads} per call (split by k), OpenMP dynamic(64) over queries",
               "topk_identical_up_to_ties_frac": agree / max(total, 1),
               "single_thread_qps": done1 / t_cpu1,
               "single_thread_sample": f"first {done1} benchmark queries, 1 thread",
               "cpu_model": model, "sockets": sockets, "u8_dot": simd,
               "recall_at_10": recall_of(cpu_labels) if n_rec and done >= n_rec else None,
               "notes": "the oracle's traversal with the reference's one-doc-ahead row prefetch, a flat visited "
                        "set per thread, and the AVX-512 gather dot product when the CPU has it"}
        log(f"cpu: {cpu['value']:,.0f} QPS on {threads} threads, top-k identical up to ties {agree}/{total}; "
            f"{cpu['single_thread_qps']:,.1f} QPS on 1 thread ({done1} queries); {model}, {sockets} socket(s), "
            f"{simd} u8 dot; CPU recall@10 {cpu['recall_at_10']}")

    # the roofline's bytes are the reference's: its traversal's counters from the
    # oracle over the same queries (the device re-scores a doc it already saw in
    # an earlier list; the reference's visited set skips it); queries the CPU
    # leg did not reach are scaled by the covered queries' ratio
    ref_bytes_src = "device counters (no CPU leg)"
    ref_bytes_launch = bytes_launch
    if cpu is not None and covered.any():
        ref_q = algorithmic_bytes(ostats[covered], probe_bytes=st[covered, 8])
        ratio = float(ref_q.sum() / dev_bytes[covered].sum())
        ref_bytes_launch = float(ref_q.sum() + ratio * dev_bytes[~covered].sum())
        ref_bytes_src = (f"oracle traversal counters over {int(covered.sum())} of {args.queries} queries"
                         + (f", the rest scaled by their ratio {ratio:.4f}" if (~covered).any() else ""))
        mismatch = int((ostats[covered, :6] != st[covered, :6]).any(axis=1).sum())
        ref_bytes_src += f"; queries whose first six counters differ from the device's: {mismatch}"
    achieved_dev = achieved
    achieved = ref_bytes_launch / (kernel_ms * 1e-3) / 1e9
    log(f"reference bytes {ref_bytes_launch / 1e9:.3f} GB/launch ({ref_bytes_src}) -> {achieved:.0f} GB/s, "
        f"{achieved / HBM_PEAK_GBS:.3f} of peak; device re-scored share "
        f"{1 - ref_bytes_launch / bytes_launch:.4f}")
    traffic, traffic_src = measured_traffic(f"seismic_sq_batched_{args.docs}docs/{args.queries}q")
    c1 = c1_reference_order(args, nsparse, threads, dev, stream)
    copy_gbs = copy_bandwidth(dev)
    log(f"device copy (2 GiB): {copy_gbs:,.0f} GB/s read+write; search achieves {achieved / copy_gbs:.3f} of it")
    read_gbs = read_bandwidth(nsparse, dev)
    traffic_gbs = traffic / (kernel_ms * 1e-3) / 1e9 if traffic else None
    log(f"device streaming read (4 GiB): {read_gbs:,.0f} GB/s"
        + (f"; the search's measured HBM traffic moves {traffic_gbs:,.0f} GB/s, {traffic_gbs / read_gbs:.3f} of it"
           if traffic_gbs else ""))
    result = {
        "metric": METRIC, "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"seismic_sq_batched_{args.docs}docs", "docs": args.docs, "queries": args.queries,
                   "k": "10 or 100 per query", "lambda": args.lam, "beta": args.beta, "alpha": args.alpha,
                   "cut": args.cut, "heap_factor": args.hf, "query_nnz": "log-uniform 5..300",
                   "launch": "hipGraph replay of the autotuned batched search",
                   "schedule": plan.config,
                   "autotuner_ms": [dict(c, ms=ms) for c, ms in tuning]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "bytes_per_launch": ref_bytes_launch, "bytes_source": ref_bytes_src,
                     "device_bytes_per_launch": bytes_launch, "device_achieved": achieved_dev,
                     "device_rescored_share": 1 - ref_bytes_launch / bytes_launch,
                     "traffic_over_bytes": (traffic / ref_bytes_launch) if traffic else None,
                     "kernel_ms": kernel_ms,
                     "copy_gbps": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "read_gbps": read_gbs, "traffic_gbps": traffic_gbs,
                     "traffic_frac_of_read": (traffic_gbs / read_gbs) if traffic_gbs else None},
        "cpu_baseline": cpu,
        "plain": {"qps": plain_qps if n_plain else None, "p50_ms": pct(0.5), "p99_ms": pct(0.99),
                  "queries": n_plain},
        "plain_host_arrays": {"qps": host_qps if n_plain else None, "p50_ms": pct_h(0.5), "p99_ms": pct_h(0.99),
                              "queries": n_plain},
        "batched_over_plain": qps / plain_qps if n_plain else None,
        "batched_over_cpu": (qps / cpu["value"]) if cpu else None,
        "recall_at_10": recall if n_rec else None, "recall_queries": n_rec,
        "recall_vs_uint8_exact": recall_sq if n_rec else None,
        "recall_minus_cpu_port": (recall - cpu["recall_at_10"]) if cpu and cpu.get("recall_at_10") is not None
        else None,
        "uint8_exact_vs_fp32_exact": ceiling if n_rec else None,
        "build_s": build_s,
        "build": build,
        "traversal_means": dict(zip(["lists", "terms_found", "csc_entries", "clusters_tested", "clusters_opened",
                                     "docs_read", "docs_scored", "nnz_scored", "probe_bytes"],
                                    st[:, :9].mean(axis=0).tolist())),
        "phase_us_per_query": dict(zip(tick_names, tick_means.tolist())),
    }
    if sweep:
        result["sweep"] = sweep
    if serving:
        result["serving"] = serving
    if inverted:
        result["inverted"] = inverted
    if rescored:
        result["two_stage_fp32_rescore"] = rescored
    if c1:
        result["c1_reference_order"] = c1
        build["cpu_baseline_c1"] = c1["cpu_build_baseline"]
        build["gpu_c1_s"] = c1["gpu_build_s"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
