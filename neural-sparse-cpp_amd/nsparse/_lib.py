"""ctypes binding of libnsparse_amd.so (the C ABI in include/nsparse_c.h).

The shared library is built in-tree next to this file. Importing this module
fails loudly when it is missing: there is no Python or CPU fallback for the
index.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsparse_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `make -C neural-sparse-cpp_amd` "
        "(or __graft_entry__.build()) first"
    )

# One HIP runtime per process: PyTorch-ROCm bundles its own libamdhip64.so.7,
# and the library resolves that soname to whichever copy is already loaded.
# Loading torch first makes both use torch's runtime (a second runtime in the
# same process leaves torch without devices).
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = ctypes.CDLL(LIB_PATH)

c_int, c_int64, c_float, c_size_t = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_size_t
c_void_p, c_char_p = ctypes.c_void_p, ctypes.c_char_p
P = ctypes.POINTER


class Selector(ctypes.Structure):
    pass


Selector._fields_ = [
    ("kind", c_int),
    ("n", c_int64),
    ("ids", P(ctypes.c_int32)),
    ("bitmap", P(ctypes.c_uint32)),
    ("inner", P(Selector)),
    ("is_member", c_void_p),
    ("ctx", c_void_p),
]


class SearchParams(ctypes.Structure):
    _fields_ = [
        ("kind", c_int),
        ("cut", c_int),
        ("heap_factor", c_float),
        ("vmin", c_float),
        ("vmax", c_float),
        ("selector", P(Selector)),
    ]


class SearchConfig(ctypes.Structure):
    _fields_ = [
        ("lookup", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
        ("longest_first", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("max_batch", ctypes.c_int32),
        ("n_streams", ctypes.c_int32),
        ("scratch_budget", ctypes.c_int64),
        ("k_max", ctypes.c_int32),
        ("cut", ctypes.c_int32),
        ("heap_factor", c_float),
    ]


class SynthSpec(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("dim", c_int),
        ("zipf_s", ctypes.c_double),
        ("nnz_mode", c_int),
        ("nnz_a", ctypes.c_double),
        ("nnz_b", ctypes.c_double),
        ("nnz_min", c_int),
        ("nnz_max", c_int),
        ("w_mu", ctypes.c_double),
        ("w_sigma", ctypes.c_double),
    ]


def _sig(name, argtypes, restype=c_int):
    fn = getattr(lib, name)
    fn.argtypes = argtypes
    fn.restype = restype
    return fn


_sig("nsparse_last_error", [], c_char_p)
_sig("nsparse_index_factory", [c_int, c_char_p, P(c_void_p)])
_sig("nsparse_free", [c_void_p])
_sig("nsparse_idmap_wrap", [c_void_p, P(c_void_p)])
_sig("nsparse_index_id", [c_void_p, ctypes.c_char * 4])
_sig("nsparse_get_dimension", [c_void_p, P(c_int)])
_sig("nsparse_num_vectors", [c_void_p, P(c_int64)])
_sig("nsparse_add", [c_void_p, c_int, c_void_p, c_void_p, c_void_p])
_sig("nsparse_add_with_ids", [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p])
_sig("nsparse_build", [c_void_p])
_sig("nsparse_search", [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                        P(SearchParams)])
_sig("nsparse_search_device", [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                               P(SearchParams), c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p])
_sig("nsparse_search_device_noalloc", [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                       P(SearchParams), c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                       c_void_p])
_sig("nsparse_search_device_stats", [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                     P(SearchParams), c_void_p, c_void_p, c_void_p, c_void_p])
_sig("nsparse_device_error", [c_void_p, P(ctypes.c_int32)])
_sig("nsparse_build_stats", [c_void_p, P(ctypes.c_int64)])
_sig("nsparse_search_plan_create", [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                    P(SearchParams), c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                    P(SearchConfig), c_void_p, P(c_void_p)])
_sig("nsparse_search_plan_launch", [c_void_p, c_void_p])
_sig("nsparse_search_plan_config", [c_void_p, P(SearchConfig), c_int, P(SearchConfig), P(c_float), P(c_int)])
_sig("nsparse_search_plan_free", [c_void_p])
_sig("nsparse_server_create", [c_void_p, P(ServerConfig), P(c_void_p)])
_sig("nsparse_server_add_bitmap", [c_void_p, c_void_p, c_int64, P(ctypes.c_int32)])
_sig("nsparse_server_submit", [c_void_p, ctypes.c_int32, c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32,
                               c_int64, P(c_int64)])
_sig("nsparse_server_wait", [c_void_p, c_int64, c_void_p, c_void_p, P(ctypes.c_double)])
_sig("nsparse_server_now_ns", [c_void_p, P(c_int64)])
_sig("nsparse_server_replay", [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p])
_sig("nsparse_server_stats", [c_void_p, c_void_p])
_sig("nsparse_server_free", [c_void_p])
_sig("nsparse_exact_scores_device", [c_int64, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p])
_sig("nsparse_rescore_device", [c_int, c_int64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int64, c_void_p,
                                c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p])
_sig("nsparse_write_index", [c_void_p, c_char_p])
_sig("nsparse_write_index_buffer", [c_void_p, P(P(ctypes.c_uint8)), P(c_size_t)])
_sig("nsparse_buffer_free", [c_void_p], None)
_sig("nsparse_read_index", [c_char_p, c_int, P(c_void_p)])
_sig("nsparse_read_index_buffer", [c_void_p, c_size_t, c_int, P(c_void_p)])
_sig("nsparse_read_csr", [c_void_p, c_char_p, c_int])
_sig("nsparse_get_vectors", [c_void_p, P(c_int64), P(c_int64), P(c_int), c_void_p, c_void_p, c_void_p])
_sig("nsparse_synth_row_nnz", [P(SynthSpec), c_int64, c_int64, c_void_p, c_int])
_sig("nsparse_synth_fill", [P(SynthSpec), c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int])
_sig("nsparse_testing_cluster_order", [c_void_p, c_int, c_void_p])
_sig("nsparse_testing_top_k_tokens", [c_void_p, c_void_p, c_int, c_int, c_void_p, P(c_int)])
_sig("nsparse_testing_device_cluster_order", [c_void_p, c_int, c_int, c_void_p])
_sig("nsparse_testing_device_top_k_tokens", [c_void_p, c_void_p, c_int, c_int, c_void_p, P(c_int)])
_sig("nsparse_testing_read_probe", [c_void_p, c_size_t, c_void_p, c_void_p])

# nsparse_status -> the exception the reference's SWIG layer raises
# (nsparse/python/swignsparse.swig:47-59).
_ERRORS = {1: ValueError, 2: RuntimeError, 3: IndexError, 4: MemoryError, 5: RuntimeError, 6: RuntimeError}


def check(status):
    if status != 0:
        msg = lib.nsparse_last_error().decode("utf-8", "replace")
        raise _ERRORS.get(status, RuntimeError)(msg)
    return status
