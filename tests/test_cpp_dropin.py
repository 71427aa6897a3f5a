"""The C++ drop-in headers (include/nsparse/): a caller written against the
reference's nsparse::Index API (tests/cpp/dropin_check.cpp) compiles unchanged
against them, links to the device library, and gets the oracle's results.

The CPU test compiles the caller and runs the calls that need no device
(factory classes, fourccs, dimensions, host-side adds, the reference's
exception types). The GPU tests build an index through index_factory / add /
build / write_index and search it through read_index with every selector
family, including IDSelector subclasses the library cannot enumerate (passed
as is_member callbacks); each result file must equal the oracle's result bit
for bit."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP, "_build", "dropin_check")
DIM = 30522


@pytest.fixture(scope="module")
def exe(nsparse):
    subprocess.run(["make", "-s", "-C", CPP], check=True)
    return EXE


def run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def write_csr(path, ip, ix, v):
    ip = np.ascontiguousarray(ip, np.int32)
    with open(path, "wb") as f:
        f.write(np.int32(len(ip) - 1).tobytes())
        f.write(np.int64(ip[-1]).tobytes())
        f.write(ip.tobytes())
        f.write(np.ascontiguousarray(ix, np.uint16).tobytes())
        f.write(np.ascontiguousarray(v, np.float32).tobytes())


def write_ids(path, ids):
    ids = np.ascontiguousarray(ids, np.int32)
    with open(path, "wb") as f:
        f.write(np.int32(len(ids)).tobytes())
        f.write(ids.tobytes())


def read_results(path, n, k):
    raw = np.fromfile(path, np.uint8)
    assert raw.size == 8 * n * k
    return raw[: 4 * n * k].view(np.float32).reshape(n, k), raw[4 * n * k:].view(np.int32).reshape(n, k)


def bitmap(member):
    words = np.zeros((len(member) + 31) // 32, np.uint32)
    for i in np.flatnonzero(member):
        words[i >> 5] |= np.uint32(1 << (i & 31))
    return words


def assert_same(got, want):
    assert (got[1] == want[1]).all(), np.argwhere(got[1] != want[1])[:5]
    assert (got[0] == want[0]).all(), np.argwhere(got[0] != want[0])[:5]


def test_headers_compile_and_host_calls(exe):
    assert run(exe, "api").strip() == "api ok"


@pytest.mark.gpu
def test_build_write_read_search(exe, tmp_path):
    from nsparse import synth

    n_docs, n_q, k = 4000, 48, 10
    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=0), n_docs)
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=1), n_q)
    vmax = np.float32(v.max())
    desc = f"seismic_sq,quantizer=8bit|vmin=0|vmax={vmax:.9g}|lambda=200|beta=20|alpha=0.4|seed=0"
    docs_bin, q_bin, ids_bin = tmp_path / "docs.bin", tmp_path / "q.bin", tmp_path / "ids.bin"
    write_csr(docs_bin, ip, ix, v)
    write_csr(q_bin, qip, qix, qv)
    allowed = np.sort(np.random.default_rng(3).choice(n_docs, n_docs // 2, replace=False)).astype(np.int32)
    write_ids(ids_bin, allowed)
    index_file = tmp_path / "sq.index"
    print("dropin build", desc, flush=True)
    run(exe, "build", DIM, desc, docs_bin, index_file)

    o = O.OracleIndex(DIM, 1, 200, 20, 0.4, 0, 0.0, vmax, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    assert index_file.read_bytes() == o.serialize(), "device build differs from the oracle"

    print("dropin search", flush=True)
    prefix = str(tmp_path / "r_")
    run(exe, "search", index_file, q_bin, ids_bin, f"{vmax:.9g}", k, prefix)
    ids = np.arange(n_docs)
    few = np.array([allowed[5], allowed[1], allowed[5], allowed[9]], np.int32)
    cases = {
        "plain": dict(cut=10, hf=0.8),
        "plain_copy": dict(cut=10, hf=0.8),
        "defaults": dict(cut=10, hf=1.0),
        "sq_range": dict(cut=10, hf=0.8, sq_range=(0.0, np.float32(2.0) * vmax)),
        "set": dict(cut=10, hf=0.8, selector=("set", allowed)),
        "not_set": dict(cut=10, hf=0.8, selector=("set", allowed), negate=True),
        "custom": dict(cut=10, hf=0.8, selector=("bitmap", bitmap(ids % 3 != 0), n_docs)),
        "not_custom": dict(cut=10, hf=0.8, selector=("bitmap", bitmap(ids % 3 == 0), n_docs)),
        "few": dict(cut=10, hf=0.8, selector=("array", few)),
    }
    for name, kw in cases.items():
        got = read_results(prefix + name + ".bin", n_q, k)
        want = o.search(qip, qix, qv, k, **kw)
        assert_same(got, want)
    # the filters did filter
    plain = read_results(prefix + "plain.bin", n_q, k)[1]
    custom = read_results(prefix + "custom.bin", n_q, k)[1]
    assert (plain[plain >= 0] % 3 == 0).any() and not (custom[custom >= 0] % 3 == 0).any()


@pytest.mark.gpu
def test_idmap_with_external_ids(exe, tmp_path):
    from nsparse import synth

    n_docs, n_q, k = 3000, 30, 10
    ip, ix, v = synth.make_corpus(synth.doc_spec(seed=3), n_docs)
    qip, qix, qv = synth.make_corpus(synth.query_spec(seed=4), n_q)
    ext = np.random.default_rng(5).permutation(n_docs).astype(np.int32) * 7 + 100
    sel = np.sort(np.random.default_rng(6).choice(ext, n_docs // 3, replace=False)).astype(np.int32)
    files = {name: tmp_path / f"{name}.bin" for name in ("docs", "ext", "q", "sel")}
    write_csr(files["docs"], ip, ix, v)
    write_ids(files["ext"], ext)
    write_csr(files["q"], qip, qix, qv)
    write_ids(files["sel"], sel)
    desc = "seismic,lambda=200|beta=20|alpha=0.4|seed=0"
    prefix = str(tmp_path / "m_")
    print("dropin idmap", desc, flush=True)
    run(exe, "idmap", DIM, desc, files["docs"], files["ext"], files["q"], files["sel"], k, prefix)
    o = O.OracleIndex(DIM, 0, 200, 20, 0.4, 0, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    nbits = int(ext.max()) + 1
    cases = {
        "plain": None,
        "set": ("set", sel),
        "custom": ("bitmap", bitmap(np.arange(nbits) % 3 != 0), nbits),
    }
    for name, selector in cases.items():
        got = read_results(prefix + name + ".bin", n_q, k)
        want = o.search_idmap(ext, qip, qix, qv, k, cut=10, hf=0.9, selector=selector)
        assert_same(got, want)
