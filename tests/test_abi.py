"""C-ABI surface on the CPU: the library loads, exports every symbol that
include/*.h declares, keeps the reference's error behaviour, and round-trips
the reference's serialized layout byte for byte. No kernel runs here."""
import ctypes
import os
import re

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(nsparse_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(nsparse):
    lib = nsparse._lib.lib
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_factory_strings_and_errors(nsparse):
    ix = nsparse.index_factory(30522, "seismic,lambda=2000|beta=200|alpha=0.4|seed=0|centroid_fraction=0.1")
    assert ix.id() == "SEIS" and ix.get_dimension() == 30522 and ix.num_vectors() == 0
    sq = nsparse.index_factory(100, "seismic_sq,quantizer=8bit|vmin=0.5|vmax=3|lambda=10|beta=2|alpha=0.4")
    assert sq.id() == "SESQ"
    im = nsparse.index_factory(100, "idmap,seismic_sq,quantizer=8bit|vmin=0.0|vmax=3.0|lambda=10|beta=2")
    assert im.id() == "IDMP"
    with pytest.raises(ValueError):
        nsparse.index_factory(10, "nope")
    with pytest.raises(ValueError):
        nsparse.index_factory(10, "")
    with pytest.raises(ValueError):
        nsparse.index_factory(10, "idmap")
    with pytest.raises(ValueError):
        nsparse.index_factory(10, "seismic_sq,vmin=2|vmax=1")
    with pytest.raises(ValueError):
        nsparse.index_factory(10, "seismic,lambda=abc")
    assert nsparse.index_factory(10, " inverted ").id() == "INVT"
    assert nsparse.index_factory(10, "idmap,inverted").id() == "IDMP"


def test_brutal_lifecycle(nsparse):
    """brutal (index_factory.cpp:95-97): created with surrounding whitespace,
    no build step (Index::build throws "not implemented"), not serializable
    (python_tests/test_error_paths.py:62-74); add works without a GPU."""
    ix = nsparse.index_factory(50, "  brutal  ")
    assert ix.id() == "BRUT" and ix.get_dimension() == 50 and ix.num_vectors() == 0
    ip = np.array([0, 2, 3], np.int32)
    ix.add(2, ip, np.array([1, 7, 3], np.uint16), np.array([0.5, 1.0, 2.0], np.float32))
    assert ix.num_vectors() == 2
    with pytest.raises(RuntimeError, match="not implemented"):
        ix.build()
    with pytest.raises(RuntimeError, match="serialization"):
        nsparse.serialize_index(ix)
    im = nsparse.index_factory(50, "idmap,brutal")
    assert im.id() == "IDMP"
    with pytest.raises(ValueError):
        ix.add(1, np.array([0, 1], np.int32), np.array([50], np.uint16), np.array([1.0], np.float32))


def test_inverted_lifecycle_and_layout(nsparse):
    """inverted (inverted_index.cpp): num_vectors counts every add, empty rows
    included; an unbuilt index writes nothing into the outputs and serializes
    with no lists; a built INVT file (here the oracle's, whose lists are pinned
    to the reference's own list-building code) reads back and re-writes byte for
    byte; bad term ids are rejected by build(); a file with another element
    width, or a list out of doc order, is rejected. No GPU needed."""
    from _oracle import OracleInverted

    ix = nsparse.index_factory(9, "inverted")
    ip = np.array([0, 1, 1, 3], np.int32)
    ix.add(3, ip, np.array([3, 0, 8], np.uint16), np.array([1.0, 0.5, 2.0], np.float32))
    assert ix.num_vectors() == 3 and ix.get_dimension() == 9
    d, lab = ix.search(1, np.array([0, 1], np.int32), np.array([3], np.uint16), np.array([1.0], np.float32), 4)
    assert (lab == -1).all() and (d == -1.0).all()
    o = OracleInverted(9)
    o.add(ip, np.array([3, 0, 8], np.uint16), np.array([1.0, 0.5, 2.0], np.float32))
    assert nsparse.serialize_index(ix) == o.serialize()
    o.build()
    built = o.serialize()
    back = nsparse.read_index_bytes(built)
    assert back.id() == "INVT" and back.num_vectors() == 3 and back.get_dimension() == 9
    assert nsparse.serialize_index(back) == built
    im = nsparse.index_factory(9, "idmap,inverted")
    im.add_with_ids(3, ip, np.array([3, 0, 8], np.uint16), np.array([1.0, 0.5, 2.0], np.float32),
                    np.array([70, 71, 72], np.int32))
    assert im.num_vectors() == 3
    bad = nsparse.index_factory(9, "inverted")
    bad.add(1, np.array([0, 1], np.int32), np.array([9], np.uint16), np.array([1.0], np.float32))
    with pytest.raises(ValueError, match="term_id out of range"):
        bad.build()
    with pytest.raises(RuntimeError, match="no vectors"):
        nsparse.index_factory(9, "inverted").build()
    wide = bytearray(built)
    wide[24:32] = np.uint64(8).tobytes()  # element width after fourcc, dim, num_vectors, n_terms
    with pytest.raises(RuntimeError, match="element width"):
        nsparse.read_index_bytes(bytes(wide))
    o2 = OracleInverted(2)
    o2.add(np.array([0, 1, 2], np.int32), np.array([0, 0], np.uint16), np.array([1.0, 2.0], np.float32))
    o2.build()
    raw = bytearray(o2.serialize())
    # list 0 holds docs [0, 1] right after its length: swap them
    at = 4 + 4 + 8 * 3 + 8
    raw[at:at + 8] = np.array([1, 0], np.int32).tobytes()
    with pytest.raises(RuntimeError, match="ascending doc order"):
        nsparse.read_index_bytes(bytes(raw))


def test_class_constructors(nsparse):
    """The SWIG module's class constructors (swignsparse.swig:149-176) with the
    classes' own defaults; IDMapIndex(index) takes ownership of its delegate."""
    assert nsparse.SeismicIndex(100).id() == "SEIS"
    p = nsparse.SeismicClusterParameters(10, 2, 0.4, 0)
    assert p.lambda_ == 10 and p.beta == 2 and p.seed == 0
    assert nsparse.SeismicIndex(100, p).get_dimension() == 100
    assert nsparse.SeismicScalarQuantizedIndex(100).id() == "SESQ"
    sq = nsparse.SeismicScalarQuantizedIndex(nsparse.QuantizerType.QT_8bit, 0.5, 3.0, p, 50)
    assert sq.id() == "SESQ" and sq.get_dimension() == 50
    q16 = nsparse.SeismicScalarQuantizedIndex(nsparse.QuantizerType.QT_16bit, 0.0, 1.0, p, 10)
    assert q16.id() == "SESQ" and q16.get_dimension() == 10
    with pytest.raises(TypeError):
        nsparse.SeismicScalarQuantizedIndex(1, 2)
    assert nsparse.BrutalIndex(5).id() == "BRUT" and nsparse.InvertedIndex(9).id() == "INVT"
    inner = nsparse.InvertedIndex(9)
    outer = nsparse.IDMapIndex(inner)
    assert outer.id() == "IDMP" and outer.get_dimension() == 9
    outer.add_with_ids(1, np.array([0, 1], np.int32), np.array([3], np.uint16), np.array([1.0], np.float32),
                       np.array([77], np.int32))
    assert outer.num_vectors() == 1
    with pytest.raises(ValueError, match="handle is null"):
        inner.num_vectors()
    assert (nsparse.Residency_kInMemory, nsparse.Residency_kMmap) == (0, 1)
    empty = nsparse.InvertedIndex(9)  # never built: search_c / search_with_params leave the padding
    q = (1, np.array([0, 1], np.int32), np.array([3], np.uint16), np.array([1.0], np.float32), 2)
    assert (empty.search_c(*q)[1] == -1).all()
    assert (empty.search_with_params(*q, nsparse.SearchParameters())[1] == -1).all()


def test_truncated_inverted_files_are_rejected(nsparse):
    """Every proper prefix of an INVT file is rejected with RuntimeError; the
    whole file reads back (io errors are runtime errors in the reference too)."""
    from _oracle import OracleInverted

    o = OracleInverted(6)
    o.add(np.array([0, 2, 3, 5], np.int32), np.array([1, 4, 0, 1, 5], np.uint16),
          np.array([0.5, 1.0, 2.0, 0.25, 3.0], np.float32))
    o.build()
    raw = o.serialize()
    for cut in range(len(raw)):
        with pytest.raises(RuntimeError):
            nsparse.read_index_bytes(raw[:cut])
    assert nsparse.serialize_index(nsparse.read_index_bytes(raw)) == raw


def test_argument_checks(nsparse):
    ix = nsparse.index_factory(10, "seismic,lambda=10|beta=2")
    i32, u16, f32 = (lambda *a: np.array(a, np.int32)), (lambda *a: np.array(a, np.uint16)), \
        (lambda *a: np.array(a, np.float32))
    with pytest.raises(ValueError):
        ix.add(0, i32(0), u16(), f32())
    with pytest.raises(ValueError):
        ix.add(1, i32(0, 1), u16(10), f32(1.0))  # term id >= dim
    # add takes exactly-typed arrays, as the SWIG typemaps do (nsparse_typemaps.i);
    # the reference fails cleanly only for values and crashes for the other two
    with pytest.raises(TypeError, match="float32 array for values"):
        ix.add(1, i32(0, 1), u16(1), np.array([1.0]))
    with pytest.raises(TypeError, match="int32 array for indptr"):
        ix.add(1, np.array([0, 1]), u16(1), f32(1.0))
    with pytest.raises(TypeError, match="uint16 array for indices"):
        ix.add(1, i32(0, 1), [1], f32(1.0))
    with pytest.raises(TypeError, match="integer"):
        ix.add(1.0, i32(0, 1), u16(1), f32(1.0))
    with pytest.raises(RuntimeError):
        ix.add_with_ids(1, [0, 1], [1], [1.0], [5])
    with pytest.raises(ValueError):
        ix.search(0, [0], [], [], 3)
    with pytest.raises(ValueError):
        ix.search(1, [0, 1], [1], [1.0], 0)
    # no vectors: the reference returns without touching the outputs
    d, lab = ix.search(1, [0, 1], [1], [1.0], 3)
    assert d.shape == (1, 3)


def test_layout_round_trip_is_byte_identical(nsparse):
    """An oracle-built index (reference layout) read by the library and
    written back comes out byte for byte the same (host-side only)."""
    rng = np.random.default_rng(2)
    n, dim = 600, 300
    ip = np.arange(0, 12 * n + 1, 12, dtype=np.int32)
    ix = np.concatenate([np.sort(rng.choice(dim, 12, replace=False)) for _ in range(n)]).astype(np.uint16)
    v = rng.lognormal(0, 0.7, 12 * n).astype(np.float32)
    for sq in (0, 1, 2):  # FP32, 8-bit codes, 16-bit codes
        o = O.OracleIndex(dim, sq, 50, 5, 0.4, 1, 0.0, float(v.max()), tie_mode=1)
        o.add(ip, ix, v)
        o.build()
        blob = o.serialize()
        got = nsparse.read_index_bytes(blob)
        assert got.id() == ("SESQ" if sq else "SEIS")
        assert got.num_vectors() == n
        assert nsparse.serialize_index(got) == blob
    # a 16-bit file whose element size says 8-bit is corrupt
    o = O.OracleIndex(dim, 2, 50, 5, 0.4, 1, 0.0, float(v.max()), tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    bad = bytearray(o.serialize())
    bad[8] = 0  # QuantizerType byte after fourcc + dim: QT_8bit with 2-byte values
    with pytest.raises(RuntimeError, match="element size"):
        nsparse.read_index_bytes(bytes(bad))


def test_corrupt_files_are_rejected(nsparse):
    rng = np.random.default_rng(3)
    ip = np.arange(0, 5 * 100 + 1, 5, dtype=np.int32)
    ix = rng.integers(0, 50, 500).astype(np.uint16)
    ix = np.concatenate([np.sort(rng.choice(50, 5, replace=False)) for _ in range(100)]).astype(np.uint16)
    v = rng.random(500).astype(np.float32)
    o = O.OracleIndex(50, 0, 20, 3, 0.4, 1, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    blob = bytearray(o.serialize())
    with pytest.raises(RuntimeError):
        nsparse.read_index_bytes(bytes(blob[: len(blob) // 2]))
    bad = bytearray(blob)
    bad[0:4] = b"XXXX"
    with pytest.raises(RuntimeError):
        nsparse.read_index_bytes(bytes(bad))


def test_file_round_trip_and_mmap(nsparse, tmp_path):
    rng = np.random.default_rng(4)
    ip = np.arange(0, 6 * 200 + 1, 6, dtype=np.int32)
    ix = np.concatenate([np.sort(rng.choice(80, 6, replace=False)) for _ in range(200)]).astype(np.uint16)
    v = rng.random(1200).astype(np.float32)
    o = O.OracleIndex(80, 1, 30, 4, 0.4, 2, 0.0, 1.0, tie_mode=1)
    o.add(ip, ix, v)
    o.build()
    p = tmp_path / "x.idx"
    p.write_bytes(o.serialize())
    for flag in (0, nsparse.kUseMmap):
        got = nsparse.read_index(str(p), flag)
        out = tmp_path / f"y{flag}.idx"
        nsparse.write_index(got, str(out))
        assert out.read_bytes() == p.read_bytes()


def test_idmap_layout(nsparse):
    """IDMP header + i2e map + nested delegate (id_map_index.cpp:76-88)."""
    im = nsparse.index_factory(20, "idmap,seismic,lambda=10|beta=2|seed=1")
    im.add_with_ids(2, [0, 1, 2], [1, 3], [0.5, 0.25], [7, 9])
    blob = nsparse.serialize_index(im)
    assert blob[:4] == b"IDMP"
    assert int.from_bytes(blob[8:16], "little") == 2
    assert np.frombuffer(blob[16:24], np.int32).tolist() == [7, 9]
    assert blob[24:28] == b"SEIS"
    back = nsparse.read_index_bytes(blob)
    assert back.id() == "IDMP" and back.num_vectors() == 2
    assert nsparse.serialize_index(back) == blob


def test_read_csr(nsparse, tmp_path):
    p = tmp_path / "m.csr"
    ip = np.array([0, 2, 3], np.int64)
    ix = np.array([1, 4, 2], np.int32)
    v = np.array([0.5, 1.0, 2.0], np.float32)
    p.write_bytes(np.array([2, 5, 3], np.int64).tobytes() + ip.tobytes() + ix.tobytes() + v.tobytes())
    idx = nsparse.index_factory(5, "seismic,lambda=10|beta=2")
    idx.read_csr(str(p))
    gip, gix, gv = idx.get_vectors()
    assert gip.tolist() == [0, 2, 3] and gix.tolist() == [1, 4, 2]
    assert np.frombuffer(gv.tobytes(), np.float32).tolist() == [0.5, 1.0, 2.0]
    with pytest.raises(ValueError):
        nsparse.index_factory(3, "seismic").read_csr(str(p))
