"""Index, search parameters and selectors over the C ABI.

Argument handling follows the reference's class_wrappers.py:22-58 (inputs are
converted with np.ascontiguousarray to int32 / uint16 / float32) and its
typemaps (outputs are freshly allocated n x k arrays).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import P, SearchParams, Selector, check, lib

kUseMmap = 1  # IndexIoFlag::kUseMmap (nsparse/io/index_io.h:17-22)

_ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731


# ---------------------------------------------------------------- selectors
class _SelectorBase:
    """IDSelector (nsparse/id_selector.h:22-86); holds its C struct and buffers."""

    def _struct(self):
        raise NotImplementedError


class _EnumerableSelector(_SelectorBase):
    _kind = None

    def __init__(self, *args):
        # SWIG exposes both (ids_array) and (n, ids_array) forms.
        ids = args[-1]
        self.ids = np.ascontiguousarray(ids, dtype=np.int32)
        if len(args) == 2:
            self.ids = self.ids[: int(args[0])]
        self._c = Selector(self._kind, len(self.ids), self.ids.ctypes.data_as(P(ctypes.c_int32)), None, None)

    def _struct(self):
        return self._c


class SetIDSelector(_EnumerableSelector):
    _kind = 1


class ArrayIDSelector(_EnumerableSelector):
    _kind = 0


class NotIDSelector(_SelectorBase):
    """Non-owning like the reference: keep `inner` alive for the search."""

    def __init__(self, inner):
        self.inner = inner
        self._c = Selector(2, 0, None, None, ctypes.pointer(inner._struct()))

    def _struct(self):
        return self._c


class BitmapIDSelector(_SelectorBase):
    """Membership bitmap over doc ids (bit i of word i // 32). Not enumerable.

    Accepts a boolean mask of length n, or uint32 words plus the bit count.
    """

    def __init__(self, mask_or_words, n_bits=None):
        arr = np.asarray(mask_or_words)
        if arr.dtype == np.bool_:
            n_bits = len(arr)
            padded = np.zeros(((n_bits + 31) // 32) * 32, dtype=np.uint8)
            padded[:n_bits] = arr
            arr = np.packbits(padded.reshape(-1, 32)[:, ::-1].reshape(-1, 8), axis=1, bitorder="big")
            arr = arr.reshape(-1, 4)[:, ::-1].copy().view(np.uint32).reshape(-1)
        self.words = np.ascontiguousarray(arr, dtype=np.uint32)
        self.n_bits = int(n_bits if n_bits is not None else 32 * len(self.words))
        self._c = Selector(3, self.n_bits, None, self.words.ctypes.data_as(P(ctypes.c_uint32)), None)

    def _struct(self):
        return self._c


# ---------------------------------------------------------------- parameters
class SearchParameters:
    """SearchParameters (nsparse/index.h:21-28): an optional, non-owning selector."""

    _kind = 0

    def __init__(self):
        self.cut = 10
        self.heap_factor = 1.0
        self.vmin = 0.0
        self.vmax = 1.0
        self._selector = None

    def set_id_selector(self, selector):
        self._selector = selector

    def get_id_selector(self):
        return self._selector

    def _struct(self):
        sel = ctypes.pointer(self._selector._struct()) if self._selector is not None else None
        return SearchParams(self._kind, int(self.cut), float(self.heap_factor), float(self.vmin),
                            float(self.vmax), sel)


class SeismicSearchParameters(SearchParameters):
    """SeismicSearchParameters(cut=10, heap_factor=1.0) (nsparse/seismic_index.h:26-32)."""

    _kind = 1

    def __init__(self, cut=10, heap_factor=1.0):
        super().__init__()
        self.cut = cut
        self.heap_factor = heap_factor


class SeismicSQSearchParameters(SeismicSearchParameters):
    """SeismicSQSearchParameters(vmin, vmax, cut, heap_factor)
    (nsparse/seismic_scalar_quantized_index.h:27-33)."""

    _kind = 2

    def __init__(self, vmin, vmax, cut=10, heap_factor=1.0):
        super().__init__(cut, heap_factor)
        self.vmin = vmin
        self.vmax = vmax


# ---------------------------------------------------------------- index
class Index:
    """Owning handle of an index created by index_factory / read_index."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.nsparse_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def id(self):
        buf = (ctypes.c_char * 4)()
        check(lib.nsparse_index_id(self._h, buf))
        return buf.raw.decode()

    def get_dimension(self):
        v = ctypes.c_int()
        check(lib.nsparse_get_dimension(self._h, ctypes.byref(v)))
        return v.value

    def num_vectors(self):
        v = ctypes.c_int64()
        check(lib.nsparse_num_vectors(self._h, ctypes.byref(v)))
        return v.value

    @staticmethod
    def _csr(indptr, indices, values):
        return (np.ascontiguousarray(indptr, dtype=np.int32), np.ascontiguousarray(indices, dtype=np.uint16),
                np.ascontiguousarray(values, dtype=np.float32))

    @staticmethod
    def _typed(arr, dtype, name):
        # the SWIG typemaps of Index.add take only arrays of the exact type
        # (nsparse_typemaps.i): the same TypeError here, for every argument
        if not isinstance(arr, np.ndarray) or arr.dtype != dtype:
            raise TypeError(f"Expected {np.dtype(dtype).name} array for {name}")
        return np.ascontiguousarray(arr)

    def add(self, n, indptr, indices, values):
        """Index.add: like the SWIG method (unlike search and add_with_ids,
        which convert), it takes int32 / uint16 / float32 arrays only."""
        if isinstance(n, bool) or not isinstance(n, (int, np.integer)):
            raise TypeError("Expected an integer for idx_t parameter")
        ip = self._typed(indptr, np.int32, "indptr")
        ix = self._typed(indices, np.uint16, "indices")
        v = self._typed(values, np.float32, "values")
        check(lib.nsparse_add(self._h, int(n), _ptr(ip), _ptr(ix), _ptr(v)))

    def add_with_ids(self, n, indptr, indices, values, ids):
        ip, ix, v = self._csr(indptr, indices, values)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        check(lib.nsparse_add_with_ids(self._h, int(n), _ptr(ip), _ptr(ix), _ptr(v), _ptr(ids)))

    def build(self):
        check(lib.nsparse_build(self._h))

    def search(self, n, indptr, indices, values, k, params=None):
        """Returns (distances float32 n x k, labels int32 n x k), best first."""
        ip, ix, v = self._csr(indptr, indices, values)
        n, k = int(n), int(k)
        # an index with nothing to search writes nothing: the padding stays
        d = np.full((max(n, 0), max(k, 0)), -1.0, dtype=np.float32)
        lab = np.full((max(n, 0), max(k, 0)), -1, dtype=np.int32)
        ps = ctypes.byref(params._struct()) if params is not None else None
        check(lib.nsparse_search(self._h, n, _ptr(ip), _ptr(ix), _ptr(v), k, _ptr(d), _ptr(lab), ps))
        return d, lab

    def search_device(self, indptr, indices, values, k_max, ks=None, params=None, out_d=None, out_l=None,
                      stream=None, stats=None, selector_of_query=None, bitmaps=None):
        """Batched search over torch tensors already on the GPU (HBM-resident
        queries); returns (distances, labels) tensors of shape n x k_max."""
        import torch

        n = indptr.numel() - 1
        dev = indptr.device
        if out_d is None:
            out_d = torch.empty((n, k_max), dtype=torch.float32, device=dev)
        if out_l is None:
            out_l = torch.empty((n, k_max), dtype=torch.int32, device=dev)
        ps = ctypes.byref(params._struct()) if params is not None else None
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        ksp = ctypes.c_void_p(ks.data_ptr()) if ks is not None else None
        if stats is not None:
            check(lib.nsparse_search_device_stats(self._h, n, indptr.data_ptr(), indices.data_ptr(),
                                                  values.data_ptr(), ksp, int(k_max), ps, out_d.data_ptr(),
                                                  out_l.data_ptr(), stats.data_ptr(), s))
        else:
            sel = ctypes.c_void_p(selector_of_query.data_ptr()) if selector_of_query is not None else None
            bm = ctypes.c_void_p(bitmaps.data_ptr()) if bitmaps is not None else None
            nb = int(bitmaps.numel()) if bitmaps is not None else 0
            check(lib.nsparse_search_device(self._h, n, indptr.data_ptr(), indices.data_ptr(), values.data_ptr(),
                                            ksp, int(k_max), ps, sel, bm, nb, out_d.data_ptr(), out_l.data_ptr(),
                                            s))
        return out_d, out_l

    def search_device_noalloc(self, indptr, indices, values, k_max, out_d, out_l, error, ks=None, params=None,
                              stream=None, selector_of_query=None, bitmaps=None):
        """search_device without any device allocation (nsparse_search_device_noalloc):
        the caller passes the outputs and a one-element int32 device tensor for
        the error code; the index must already be resident on the device."""
        n = indptr.numel() - 1
        ps = ctypes.byref(params._struct()) if params is not None else None
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        ksp = ctypes.c_void_p(ks.data_ptr()) if ks is not None else None
        sel = ctypes.c_void_p(selector_of_query.data_ptr()) if selector_of_query is not None else None
        bm = ctypes.c_void_p(bitmaps.data_ptr()) if bitmaps is not None else None
        nb = int(bitmaps.numel()) if bitmaps is not None else 0
        check(lib.nsparse_search_device_noalloc(self._h, n, indptr.data_ptr(), indices.data_ptr(),
                                                values.data_ptr(), ksp, int(k_max), ps, sel, bm, nb,
                                                out_d.data_ptr(), out_l.data_ptr(), error.data_ptr(), s))
        return out_d, out_l

    def search_plan(self, indptr, indices, values, k_max, ks=None, params=None, out_d=None, out_l=None,
                    config=None, stream=None, selector_of_query=None, bitmaps=None):
        """A SearchPlan: the batched search over these GPU tensors captured in a
        hipGraph (nsparse_search_plan_create). config None: autotuned on this
        batch; else a dict with lookup / block / persistent / longest_first."""
        import torch

        n = indptr.numel() - 1
        dev = indptr.device
        if out_d is None:
            out_d = torch.empty((n, k_max), dtype=torch.float32, device=dev)
        if out_l is None:
            out_l = torch.empty((n, k_max), dtype=torch.int32, device=dev)
        ps = ctypes.byref(params._struct()) if params is not None else None
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        ksp = ctypes.c_void_p(ks.data_ptr()) if ks is not None else None
        sel = ctypes.c_void_p(selector_of_query.data_ptr()) if selector_of_query is not None else None
        bm = ctypes.c_void_p(bitmaps.data_ptr()) if bitmaps is not None else None
        nb = int(bitmaps.numel()) if bitmaps is not None else 0
        cfg = None
        if config is not None:
            cfg = ctypes.byref(_lib.SearchConfig(int(config.get("lookup", 1)), int(config.get("block", 512)),
                                                 int(config.get("persistent", 0)),
                                                 int(config.get("longest_first", 0))))
        h = ctypes.c_void_p()
        check(lib.nsparse_search_plan_create(self._h, n, indptr.data_ptr(), indices.data_ptr(), values.data_ptr(),
                                             ksp, int(k_max), ps, sel, bm, nb, out_d.data_ptr(), out_l.data_ptr(),
                                             cfg, s, ctypes.byref(h)))
        # the plan keeps raw pointers: hold the tensors as long as it lives
        return SearchPlan(h.value, (indptr, indices, values, ks, out_d, out_l, selector_of_query, bitmaps, self))

    def device_error(self):
        v = ctypes.c_int32()
        check(lib.nsparse_device_error(self._h, ctypes.byref(v)))
        return v.value

    def build_stats(self):
        """What the last build() did: kept postings, their rows' non-zeros,
        clusters, summary entries, assignment (doc term, centroid posting) pairs."""
        v = (ctypes.c_int64 * 5)()
        check(lib.nsparse_build_stats(self._h, v))
        return dict(zip(["kept_postings", "kept_row_nnz", "clusters", "summary_entries", "assign_pairs"], v))

    # the SWIG extension methods the reference's wrappers build search() from
    # (swignsparse.swig, %extend nsparse::Index)
    def search_c(self, n, indptr, indices, values, k):
        return self.search(n, indptr, indices, values, k)

    def search_with_params(self, n, indptr, indices, values, k, params):
        return self.search(n, indptr, indices, values, k, params)

    def read_csr(self, path, residency=0):
        check(lib.nsparse_read_csr(self._h, str(path).encode(), int(residency)))

    def get_vectors(self):
        """Stored forward index as (indptr int32, indices uint16, values bytes)."""
        n, nnz, esz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        check(lib.nsparse_get_vectors(self._h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(esz), None, None,
                                      None))
        ip = np.empty(n.value + 1 if n.value else 0, dtype=np.int32)
        ix = np.empty(nnz.value, dtype=np.uint16)
        v = np.empty(nnz.value * esz.value, dtype=np.uint8)
        check(lib.nsparse_get_vectors(self._h, None, None, None, _ptr(ip), _ptr(ix), _ptr(v)))
        return ip, ix, v


class SearchPlan:
    """A captured batched search (see Index.search_plan)."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep
        self.out_d, self.out_l = keep[4], keep[5]

    def launch(self, stream=None):
        s = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
        check(lib.nsparse_search_plan_launch(self._h, s))
        return self.out_d, self.out_l

    @property
    def config(self):
        c = _lib.SearchConfig()
        n = ctypes.c_int()
        check(lib.nsparse_search_plan_config(self._h, ctypes.byref(c), 0, None, None, ctypes.byref(n)))
        return c.as_dict()

    @property
    def tuning(self):
        """[(config dict, best ms)] of every schedule the autotuner timed."""
        n = ctypes.c_int()
        check(lib.nsparse_search_plan_config(self._h, None, 0, None, None, ctypes.byref(n)))
        cs = (_lib.SearchConfig * max(n.value, 1))()
        ms = (ctypes.c_float * max(n.value, 1))()
        check(lib.nsparse_search_plan_config(self._h, None, n.value, cs, ms, ctypes.byref(n)))
        return [(cs[i].as_dict(), float(ms[i])) for i in range(n.value)]

    def free(self):
        if self._h:
            check(lib.nsparse_search_plan_free(self._h))
            self._h = None
            self._keep = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Continuous-batching search server over a built SEISMIC index resident in
    HBM (nsparse_server_*): requests are batched as they arrive, batches run on
    several streams within a scratch-memory budget."""

    STAT_NAMES = ("completed", "batches", "max_inflight_bytes", "max_batch", "deferred_by_budget",
                  "max_concurrent_batches")

    def __init__(self, index, max_batch=512, streams=4, scratch_budget=256 << 20, k_max=100, cut=10,
                 heap_factor=1.0):
        cfg = _lib.ServerConfig(int(max_batch), int(streams), int(scratch_budget), int(k_max), int(cut),
                                float(heap_factor))
        h = ctypes.c_void_p()
        check(lib.nsparse_server_create(index.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self._index = index
        self.k_max = int(k_max)

    def add_bitmap(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        out = ctypes.c_int32()
        check(lib.nsparse_server_add_bitmap(self._h, _ptr(w), len(w), ctypes.byref(out)))
        return out.value

    def submit(self, indices, values, k, bitmap_id=-1, arrival_ns=-1):
        ix = np.ascontiguousarray(indices, dtype=np.uint16)
        v = np.ascontiguousarray(values, dtype=np.float32)
        t = ctypes.c_int64()
        check(lib.nsparse_server_submit(self._h, len(ix), _ptr(ix), _ptr(v), int(k), int(bitmap_id),
                                        int(arrival_ns), ctypes.byref(t)))
        return t.value

    def wait(self, ticket, k):
        d = np.empty(k, np.float32)
        lab = np.empty(k, np.int32)
        us = ctypes.c_double()
        check(lib.nsparse_server_wait(self._h, int(ticket), _ptr(d), _ptr(lab), ctypes.byref(us)))
        return d, lab, us.value

    def replay(self, arrival_s, indptr, indices, values, ks, bitmap_ids=None):
        """Submits query i at arrival_s[i] seconds from now; returns (distances,
        labels) of shape n x k_max and the latencies in microseconds."""
        n = len(indptr) - 1
        a = np.ascontiguousarray(arrival_s, dtype=np.float64)
        ip = np.ascontiguousarray(indptr, dtype=np.int32)
        ix = np.ascontiguousarray(indices, dtype=np.uint16)
        v = np.ascontiguousarray(values, dtype=np.float32)
        kk = np.ascontiguousarray(ks, dtype=np.int32)
        bm = np.ascontiguousarray(bitmap_ids, dtype=np.int32) if bitmap_ids is not None else None
        d = np.empty((n, self.k_max), np.float32)
        lab = np.empty((n, self.k_max), np.int32)
        lat = np.empty(n, np.float64)
        check(lib.nsparse_server_replay(self._h, n, _ptr(a), _ptr(ip), _ptr(ix), _ptr(v), _ptr(kk),
                                        _ptr(bm) if bm is not None else None, _ptr(d), _ptr(lab), _ptr(lat)))
        return d, lab, lat

    def stats(self):
        out = np.zeros(6, np.int64)
        check(lib.nsparse_server_stats(self._h, _ptr(out)))
        return dict(zip(self.STAT_NAMES, out.tolist()))

    def close(self):
        if self._h:
            check(lib.nsparse_server_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def index_factory(dim, description):
    h = ctypes.c_void_p()
    check(lib.nsparse_index_factory(int(dim), description.encode(), ctypes.byref(h)))
    return Index(h.value)


# ---------------------------------------------------------------- constructors
# The SWIG module also exposes the index classes' own constructors
# (swignsparse.swig:149-176). Their defaults are the classes' (seismic_common.h:
# 36, 50-58: lambda and beta -1 = derived from the corpus, alpha 0.4, random
# seed; scalar_quantizer.h:30: 8-bit over [0, 1]), not the factory's; each
# returns an Index built through the factory string that says the same.
# Residency (types.h:27-30), as SWIG names a scoped enum's values
Residency_kInMemory = 0
Residency_kMmap = 1


class QuantizerType:
    """utils/scalar_quantizer.h:21-24."""

    QT_8bit = 0
    QT_16bit = 1


class SeismicClusterParameters:
    """SeismicClusterParameters (seismic_common.h:38-47); `lambda` is a Python
    keyword, so the field is `lambda_`, as SWIG names it."""

    def __init__(self, lambda_=-1, beta=-1, alpha=0.4, seed=-1, inverted_list_batch_size=1):
        self.lambda_ = int(lambda_)
        self.beta = int(beta)
        self.alpha = float(alpha)
        self.seed = int(seed)
        self.inverted_list_batch_size = int(inverted_list_batch_size)

    def _desc(self):
        return f"lambda={self.lambda_}|beta={self.beta}|alpha={self.alpha:.9g}|seed={self.seed}"


def SeismicIndex(dim, parameter=None):  # noqa: N802 -- the reference's class name
    """SeismicIndex(dim) / SeismicIndex(dim, SeismicClusterParameters) (seismic_index.h:38-39)."""
    p = parameter if parameter is not None else SeismicClusterParameters()
    return index_factory(dim, "seismic," + p._desc())


def SeismicScalarQuantizedIndex(*args):  # noqa: N802
    """SeismicScalarQuantizedIndex(dim) or (quantizer_type, vmin, vmax, parameter, dim)
    (seismic_scalar_quantized_index.h:38-41)."""
    if len(args) == 1:
        qt, vmin, vmax, p, dim = QuantizerType.QT_8bit, 0.0, 1.0, SeismicClusterParameters(), args[0]
    elif len(args) == 5:
        qt, vmin, vmax, p, dim = args
    else:
        raise TypeError("SeismicScalarQuantizedIndex(dim) or (quantizer_type, vmin, vmax, parameter, dim)")
    q = "8bit" if qt == QuantizerType.QT_8bit else "16bit"
    return index_factory(dim, f"seismic_sq,quantizer={q}|vmin={float(vmin):.9g}|vmax={float(vmax):.9g}|" + p._desc())


def BrutalIndex(dim=0):  # noqa: N802
    """BrutalIndex(dim = 0) (brutal_index.h:24)."""
    return index_factory(dim, "brutal")


def InvertedIndex(dim):  # noqa: N802
    """InvertedIndex(dim) (inverted_index.h:27)."""
    return index_factory(dim, "inverted")


def IDMapIndex(index):  # noqa: N802
    """IDMapIndex(Index*) (id_map_index.h:84): takes ownership of `index`, whose
    handle holds no index afterwards."""
    h = ctypes.c_void_p()
    check(lib.nsparse_idmap_wrap(index.handle, ctypes.byref(h)))
    return Index(h.value)


def write_index(index, path):
    check(lib.nsparse_write_index(index.handle, str(path).encode()))


def serialize_index(index):
    """write_index into memory; returns bytes."""
    buf = P(ctypes.c_uint8)()
    size = ctypes.c_size_t()
    check(lib.nsparse_write_index_buffer(index.handle, ctypes.byref(buf), ctypes.byref(size)))
    try:
        out = bytearray(size.value)
        if size.value:
            ctypes.memmove((ctypes.c_char * size.value).from_buffer(out), buf, size.value)
        return bytes(out)
    finally:
        lib.nsparse_buffer_free(buf)


def read_index(path, io_flags=0):
    h = ctypes.c_void_p()
    check(lib.nsparse_read_index(str(path).encode(), int(io_flags), ctypes.byref(h)))
    return Index(h.value)


def read_index_bytes(data, io_flags=0):
    h = ctypes.c_void_p()
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    check(lib.nsparse_read_index_buffer(buf, len(data), int(io_flags), ctypes.byref(h)))
    return Index(h.value)
