"""ctypes declarations of include/qann_hip.h for tests that call the C ABI
directly (no host index objects in between)."""
import ctypes as C

import numpy as np

from conftest import ROOT

OK, EINVAL, ESTATE, ENOMEM, EDEVICE = 0, 1, 2, 3, 4
METRIC_L2, METRIC_COSINE, METRIC_IP = 0, 1, 2
IVF_FASTSCAN, IVF_PQ, IVF_SQ, IVF_FLOAT = 0, 1, 2, 3


class IvfDesc(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("nlist", C.c_uint32), ("metric", C.c_uint32), ("m", C.c_uint32),
                ("cosine_ip", C.c_uint32), ("use_precomputed", C.c_uint32),
                ("centroids", C.c_void_p), ("residual_centers", C.c_void_p), ("pq_centroids", C.c_void_p),
                ("rotation", C.c_void_p), ("storage", C.c_uint32), ("pq_k", C.c_uint32), ("sq_bits", C.c_uint32),
                ("sq_vmin", C.c_void_p), ("sq_step", C.c_void_p), ("pq_sdc", C.c_uint32)]


class SearchParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("nprobe", C.c_uint32), ("k_factor", C.c_uint32),
                ("filter_probe_scaling", C.c_uint32), ("allowed_words", C.c_void_p), ("allowed_count", C.c_uint64)]


def lib():
    h = C.CDLL(str(ROOT / "qann_amd" / "libqann_hip.so"))
    h.qann_hip_last_error.restype = C.c_char_p
    h.qann_hip_index_destroy.restype = None
    h.qann_hip_index_destroy.argtypes = [C.c_void_p]
    for name in ("qann_hip_ivf_create",):
        getattr(h, name).argtypes = [C.POINTER(IvfDesc), C.POINTER(C.c_void_p)]
    h.qann_hip_ivf_set_lists.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    h.qann_hip_refine_create.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]
    h.qann_hip_flat_create.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]
    h.qann_hip_set_ids.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    h.qann_hip_search.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(SearchParams), C.c_void_p,
                                  C.c_void_p]
    h.qann_hip_search_async.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(SearchParams), C.c_void_p,
                                        C.c_void_p, C.c_void_p]
    return h


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Handle:
    """A device index created through the C ABI; destroyed with the object."""

    def __init__(self, lib_, h, keep=()):
        self.lib, self.h, self.keep = lib_, h, list(keep)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.qann_hip_index_destroy(self.h)
            self.h = None

    def check(self, rc):
        if rc != OK:
            raise RuntimeError(f"status {rc}: {self.lib.qann_hip_last_error().decode()}")

    def set_ids(self, slots, labels=None, deleted=None):
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int64)
        dele = None if deleted is None else np.ascontiguousarray(deleted, dtype=np.uint8)
        self.check(self.lib.qann_hip_set_ids(self.h, slots, ptr(lab), ptr(dele)))

    def search(self, q, k, nprobe=1, k_factor=0, allowed_words=None, allowed_count=0, scaling=1):
        q = np.ascontiguousarray(q, dtype=np.float32)
        ids = np.empty((q.shape[0], k), np.int64)
        d = np.empty((q.shape[0], k), np.float32)
        aw = None if allowed_words is None else np.ascontiguousarray(allowed_words, dtype=np.uint64)
        p = SearchParams(k=k, nprobe=nprobe, k_factor=k_factor, filter_probe_scaling=scaling,
                         allowed_words=ptr(aw), allowed_count=allowed_count)
        self.check(self.lib.qann_hip_search(self.h, ptr(q), q.shape[0], C.byref(p), ptr(ids), ptr(d)))
        return ids, d


def float_ivf(lib_, centroids, lists, rows, metric=METRIC_L2):
    """An IVF index of float lists (storage QANN_IVF_FLOAT) through the C ABI:
    lists[l] holds the slots of list l, rows the float row of every slot."""
    cent = np.ascontiguousarray(centroids, dtype=np.float32)
    nlist, dim = cent.shape
    d = IvfDesc(dim=dim, nlist=nlist, metric=metric, m=0, centroids=ptr(cent), storage=IVF_FLOAT)
    out = C.c_void_p()
    rc = lib_.qann_hip_ivf_create(C.byref(d), C.byref(out))
    if rc != OK:
        raise RuntimeError(lib_.qann_hip_last_error().decode())
    h = Handle(lib_, out.value, keep=[cent])
    sizes = np.array([len(x) for x in lists], dtype=np.uint32)
    slots = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.uint32) for x in lists]))
    codes = np.ascontiguousarray(rows[slots.astype(np.int64)], dtype=np.float32)
    h.check(lib_.qann_hip_ivf_set_lists(h.h, ptr(sizes), ptr(slots), ptr(codes)))
    return h
