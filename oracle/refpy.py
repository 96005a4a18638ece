"""ctypes bindings for the two CPU checkers built by oracle/Makefile.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use these to check and time the product; the product path
(qann_amd) never imports this module.

* ``Oracle`` wraps liboracle.so, this project's CPU restatement of the
  reference's IVF fast-scan + refine + flat search (oracle/qann_oracle.cpp).
* ``Reference`` wraps _ref/libqann_ref.so, the reference library itself built
  from its own sources with oracle/ref_harness.cpp (only where the reference
  sources were present at build time).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libqann_ref.so"

_f32p = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u64p = C.POINTER(C.c_uint64)


def _fp(a):
    return a.ctypes.data_as(_f32p)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_i64p)


class _Lib:
    prefix = ""

    def __init__(self, path: pathlib.Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (run: make -C {HERE})")
        self.lib = C.CDLL(str(path))
        self._declare()

    def _fn(self, name, res, *args):
        f = getattr(self.lib, self.prefix + name)
        f.restype = res
        f.argtypes = list(args)
        return f

    def _check(self, rc):
        if rc != 0:
            msg = self.last_error().decode()
            raise (ValueError if rc == 1 else RuntimeError)(msg)


class _Index:
    """A loaded or built index inside one of the checker libraries."""

    def __init__(self, lib: "_Lib", handle):
        self._lib = lib
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.free(self._h)
            self._h = None

    def info(self):
        out = np.zeros(8, dtype=np.int64)
        self._lib._check(self._lib.info(self._h, _ip(out)))
        return {"kind": int(out[0]), "dim": int(out[1]), "size": int(out[2]), "slots": int(out[3]),
                "metric": int(out[4])}

    @property
    def dim(self):
        return self.info()["dim"]

    def set_nprobe(self, nprobe):
        self._lib._check(self._lib.set_nprobe(self._h, int(nprobe)))

    def set_k_factor(self, kf):
        self._lib._check(self._lib.set_k_factor(self._h, int(kf)))

    def set_filter_probe_scaling(self, on):
        self._lib._check(self._lib.set_filter_probe_scaling(self._h, int(bool(on))))

    def set_pq_distance(self, sdc):
        self._lib._check(self._lib.set_pq_distance(self._h, int(bool(sdc))))

    def batch_query(self, queries, k, allowed=None, where=None):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), dtype=np.int64)
        dists = np.empty((nq, k), dtype=np.float32)
        al = None if allowed is None else np.ascontiguousarray(allowed, dtype=np.int64)
        args = [self._h, _fp(q), nq, int(k), _ip(al), 0 if al is None else al.shape[0]]
        if isinstance(self._lib, Reference):
            args.append(None if where is None else where.encode())
        elif where is not None:
            raise NotImplementedError("where filters are evaluated by the reference only")
        self._lib._check(self._lib.batch_query(*args, _ip(ids), _fp(dists)))
        return ids, dists


class Reference(_Lib):
    prefix = "ref_"

    def __init__(self, path: pathlib.Path = REF_SO, disable_avx512: bool = True):
        # The reference picks its AVX-512 distance kernels once, at load, unless
        # QANN_DISABLE_AVX512 is set; the oracle restates the AVX2 summation order.
        if disable_avx512:
            os.environ["QANN_DISABLE_AVX512"] = "1"
        super().__init__(path)

    def _declare(self):
        self.last_error = self._fn("last_error", C.c_char_p)
        self.set_threads = self._fn("set_threads", None, C.c_uint64)
        self.load_ = self._fn("load", C.c_void_p, C.c_char_p)
        self.free = self._fn("free", None, C.c_void_p)
        self.info = self._fn("info", C.c_int, C.c_void_p, _i64p)
        self.set_nprobe = self._fn("set_nprobe", C.c_int, C.c_void_p, C.c_uint64)
        self.set_k_factor = self._fn("set_k_factor", C.c_int, C.c_void_p, C.c_uint64)
        self.set_precomputed = self._fn("set_precomputed", C.c_int, C.c_void_p, C.c_int)
        self.set_cosine_ip = self._fn("set_cosine_ip", C.c_int, C.c_void_p, C.c_int)
        self.set_pq_distance = self._fn("set_pq_distance", C.c_int, C.c_void_p, C.c_int)
        self.set_filter_probe_scaling = self._fn("set_filter_probe_scaling", C.c_int, C.c_void_p, C.c_int)
        self.batch_query = self._fn("batch_query", C.c_int, C.c_void_p, _f32p, C.c_uint64, C.c_uint64, _i64p,
                                    C.c_int64, C.c_char_p, _i64p, _f32p)
        self.build_ivf_ = self._fn("build_ivf", C.c_void_p, _f32p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                   C.c_int, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                   C.c_uint64, _i64p, C.c_int, C.c_uint64, C.c_uint64)
        self.build_flat_ = self._fn("build_flat", C.c_void_p, _f32p, C.c_uint64, C.c_uint64, C.c_int, _i64p)
        self.add_ = self._fn("add", C.c_int, C.c_void_p, _f32p, C.c_uint64, _i64p)
        self.remove_ = self._fn("remove", C.c_int, C.c_void_p, _i64p, C.c_uint64, _u64p)
        self.save_ = self._fn("save", C.c_int, C.c_void_p, C.c_char_p)
        self.quantize_lut_ = self._fn("quantize_lut", None, _f32p, C.c_uint64, _u8p, _f32p, _f32p)
        self.build_quantize_lut_ = self._fn("build_quantize_lut", None, _f32p, _f32p, C.c_float, C.c_uint64, _f32p,
                                            _u8p, _f32p, _f32p)
        self.fast_scan_block_mask_ = self._fn("fast_scan_block_mask", C.c_uint32, _u8p, _u8p, C.c_uint64,
                                              C.c_uint16, _u16p)
        self.l2_distance_ = self._fn("l2_distance", C.c_float, _f32p, _f32p, C.c_uint64)
        self.inner_product_ = self._fn("inner_product", C.c_float, _f32p, _f32p, C.c_uint64)

    def load(self, path) -> _Index:
        h = self.load_(str(path).encode())
        if not h:
            raise RuntimeError(self.last_error().decode())
        return _Index(self, h)

    def build_ivf(self, data, nlist, nprobe, *, metric=0, m=0, pq_k=16, fast_scan=True, rotate=False, sq_bits=0,
                  seed=1, k_factor=0, ids=None, clustering=0, iters_or_passes=2, train_n=0) -> _Index:
        x = np.ascontiguousarray(data, dtype=np.float32)
        idv = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        h = self.build_ivf_(_fp(x), x.shape[0], x.shape[1], nlist, nprobe, metric, m, pq_k, int(fast_scan),
                            int(rotate), sq_bits, seed, k_factor, _ip(idv), clustering, iters_or_passes, train_n)
        if not h:
            raise RuntimeError(self.last_error().decode())
        return _Index(self, h)

    def build_flat(self, data, *, metric=0, ids=None) -> _Index:
        x = np.ascontiguousarray(data, dtype=np.float32)
        idv = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        h = self.build_flat_(_fp(x), x.shape[0], x.shape[1], metric, _ip(idv))
        if not h:
            raise RuntimeError(self.last_error().decode())
        return _Index(self, h)

    def add(self, index: _Index, data, ids=None):
        x = np.ascontiguousarray(data, dtype=np.float32)
        idv = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        self._check(self.add_(index._h, _fp(x), x.shape[0], _ip(idv)))

    def remove(self, index: _Index, ids):
        idv = np.ascontiguousarray(ids, dtype=np.int64)
        out = C.c_uint64(0)
        self._check(self.remove_(index._h, _ip(idv), idv.shape[0], C.byref(out)))
        return out.value

    def save(self, index: _Index, path):
        self._check(self.save_(index._h, str(path).encode()))

    def quantize_lut(self, table, m):
        t = np.ascontiguousarray(table, dtype=np.float32)
        lut = np.zeros(m * 16, dtype=np.uint8)
        s, o = C.c_float(0), C.c_float(0)
        self.quantize_lut_(_fp(t), m, lut.ctypes.data_as(_u8p), C.byref(s), C.byref(o))
        return lut, np.float32(s.value), np.float32(o.value)

    def fast_scan_block_mask(self, lut, block, m, limit):
        lut = np.ascontiguousarray(lut, dtype=np.uint8)
        block = np.ascontiguousarray(block, dtype=np.uint8)
        out = np.zeros(32, dtype=np.uint16)
        mask = self.fast_scan_block_mask_(lut.ctypes.data_as(_u8p), block.ctypes.data_as(_u8p), m, limit,
                                          out.ctypes.data_as(_u16p))
        return mask, out

    def l2_distance(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        return np.float32(self.l2_distance_(_fp(a), _fp(b), a.shape[0]))

    def inner_product(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        return np.float32(self.inner_product_(_fp(a), _fp(b), a.shape[0]))


class Oracle(_Lib):
    prefix = "orc_"

    def __init__(self, path: pathlib.Path = ORACLE_SO):
        super().__init__(path)

    def _declare(self):
        self.last_error = self._fn("last_error", C.c_char_p)
        self.set_threads = self._fn("set_threads", None, C.c_uint64)
        self.load_ = self._fn("load", C.c_void_p, C.c_char_p)
        self.free = self._fn("free", None, C.c_void_p)
        self.info = self._fn("info", C.c_int, C.c_void_p, _i64p)
        self.set_nprobe = self._fn("set_nprobe", C.c_int, C.c_void_p, C.c_uint64)
        self.set_k_factor = self._fn("set_k_factor", C.c_int, C.c_void_p, C.c_uint64)
        self.set_filter_probe_scaling = self._fn("set_filter_probe_scaling", C.c_int, C.c_void_p, C.c_int)
        self.batch_query = self._fn("batch_query", C.c_int, C.c_void_p, _f32p, C.c_uint64, C.c_uint64, _i64p,
                                    C.c_int64, _i64p, _f32p)
        self.quantize_lut_ = self._fn("quantize_lut", None, _f32p, C.c_uint64, _u8p, _f32p, _f32p)
        self.fast_scan_block_mask_ = self._fn("fast_scan_block_mask", C.c_uint32, _u8p, _u8p, C.c_uint64,
                                              C.c_uint16, _u16p)
        self.l2_distance_ = self._fn("l2_distance", C.c_float, _f32p, _f32p, C.c_uint64)
        self.inner_product_ = self._fn("inner_product", C.c_float, _f32p, _f32p, C.c_uint64)
        self.probes_ = self._fn("probes", C.c_int, C.c_void_p, _f32p, C.c_uint64, C.POINTER(C.c_uint16), _f32p)

    def load(self, path) -> _Index:
        h = self.load_(str(path).encode())
        if not h:
            raise RuntimeError(self.last_error().decode())
        return _Index(self, h)

    quantize_lut = Reference.quantize_lut
    fast_scan_block_mask = Reference.fast_scan_block_mask
    l2_distance = Reference.l2_distance
    inner_product = Reference.inner_product

    def probes(self, index: _Index, queries, nprobe):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        lists = np.zeros((q.shape[0], nprobe), dtype=np.uint16)
        dists = np.zeros((q.shape[0], nprobe), dtype=np.float32)
        self._check(self.probes_(index._h, _fp(q), q.shape[0], lists.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 _fp(dists)))
        return lists, dists
