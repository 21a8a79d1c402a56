"""hora_amd -- MI355X (gfx950) engine for hora's distance-and-scan path.

The compute lives in libhora_mi355x.so (HIP kernels + C++ runtime behind the
C ABI of include/hora_mi355x.h); this package is the Python mirror of
hora's ANNIndex surface over that ABI.
"""
from ._lib import HoraError, build, declared_symbols, lib  # noqa: F401
from .index import (  # noqa: F401
    BruteForceIndex,
    BruteForceParams,
    HM_SEARCH_BF16_RERANK,
    HM_SEARCH_EXACT,
    HM_SEARCH_FP8_RERANK,
    IVFPQIndex,
    IVFPQParams,
    Metric,
    Node,
    PQIndex,
    PQIVFIndex,
    PQParams,
    naive_build_knn_graph,
)
