"""MI355X-native SEISMIC / SEISMIC-SQ index with the reference's Python surface.

Mirrors neural-sparse-cpp's SWIG module (nsparse/python/swignsparse.swig,
class_wrappers.py): ``index_factory``, ``Index.add / add_with_ids / build /
search``, ``SearchParameters``, ``SeismicSearchParameters``,
``SeismicSQSearchParameters``, ``SetIDSelector`` / ``ArrayIDSelector`` /
``NotIDSelector``, ``write_index`` / ``read_index`` and ``kUseMmap``. Every call
goes through the C ABI of libnsparse_amd.so; search and build run on the GPU.
"""
from .index import (  # noqa: F401
    ArrayIDSelector,
    BitmapIDSelector,
    BrutalIndex,
    IDMapIndex,
    Index,
    InvertedIndex,
    NotIDSelector,
    QuantizerType,
    Residency_kInMemory,
    Residency_kMmap,
    SeismicClusterParameters,
    SeismicIndex,
    SeismicScalarQuantizedIndex,
    SearchParameters,
    SearchPlan,
    Server,
    SeismicSearchParameters,
    SeismicSQSearchParameters,
    SetIDSelector,
    index_factory,
    kUseMmap,
    read_index,
    read_index_bytes,
    serialize_index,
    write_index,
)
from .synth import SynthSpec, make_corpus  # noqa: F401

__version__ = "0.1.0"
