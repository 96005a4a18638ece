"""qann_amd -- qann's IVF 4-bit PQ fast-scan search with exact re-ranking on
AMD MI355X (gfx950).

The API mirrors the reference's Python module ``qann`` for the dense indexes
on this path: ``FlatIndex``, ``IVFIndex`` (fast-scan PQ), ``RefineIndex``,
``Metric``, ``load`` and the ``where=`` metadata filters (``field``, ``where``,
``Expr``). Files saved by ``qann`` load here unchanged and
files saved here load in ``qann``. Every query runs on the GPU through the
HIP library ``libqann_hip.so``; there is no CPU search path, so importing
this package fails loudly when the native libraries are not built.
"""
from __future__ import annotations

import pathlib

import importlib.util

_here = pathlib.Path(__file__).resolve().parent
# torch ships its own HIP runtime under the same soname as ROCm's. Whichever
# loads first serves the process, and torch only works with its own; so when
# torch is installed it is imported first, and both share its runtime.
if importlib.util.find_spec("torch") is not None:
    import torch  # noqa: F401

try:
    from . import _qann  # noqa: F401  (loads libqann_hip.so through its rpath)
except ImportError as e:  # pragma: no cover - exercised only on a broken build
    raise ImportError(
        f"qann_amd's native extension is not built or cannot load ({e}); run `python qann_amd/build.py`"
    ) from e

from ._qann import (  # noqa: E402
    Clustering,
    DenseIndex,
    Expr,
    Field,
    FlatIndex,
    IVFIndex,
    Metric,
    PQDistance,
    RefineIndex,
    Server,
    ServeFuture,
    all_of,
    any_of,
    field,
    get_max_filter_conditions,
    get_num_threads,
    load,
    serve_load,
    set_max_filter_conditions,
    set_num_threads,
    where,
)

Index = DenseIndex
__version__ = "0.1.0"
__all__ = ["Clustering", "DenseIndex", "Index", "FlatIndex", "IVFIndex", "RefineIndex", "Metric", "PQDistance", "load", "set_num_threads",
           "get_num_threads", "native_library", "Server", "ServeFuture", "serve_load", "Expr", "Field", "field",
           "where", "any_of", "all_of", "get_max_filter_conditions", "set_max_filter_conditions"]


def native_library() -> pathlib.Path:
    """Path of the HIP library every search runs through."""
    return _here / "libqann_hip.so"
