"""qann_hip_search_noalloc (the search that allocates no device memory) through
the C ABI against qann_hip_search_async on the same device inputs: ids and
distance bits equal, on two calls in a row (the second with fewer queries),
unfiltered and filtered, IVF and RefineIndex handles; and its refusals."""
import ctypes as C

import numpy as np
import pytest

import capi
import qann_amd
from cases import synth_case

pytestmark = pytest.mark.gpu


def _lib():
    lib = capi.lib()
    lib.qann_hip_search_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(capi.SearchParams),
                                            C.c_void_p, C.c_void_p, C.c_void_p]
    lib.qann_hip_noalloc_scratch_bytes.restype = C.c_uint64
    return lib


def _run(lib, fn, handle, dq, nq, params, k, stream):
    import torch

    ids = torch.full((nq, k), -7, dtype=torch.int64, device="cuda")
    d = torch.full((nq, k), 7.0, dtype=torch.float32, device="cuda")
    rc = fn(C.c_void_p(handle), C.c_void_p(dq.data_ptr()), C.c_uint64(nq), C.byref(params), C.c_void_p(ids.data_ptr()),
            C.c_void_p(d.data_ptr()), C.c_void_p(stream))
    torch.cuda.synchronize()
    assert rc == 0, lib.qann_hip_last_error().decode()
    return ids.cpu().numpy(), d.cpu().numpy()


def _filter_words(slots, allowed_slots):
    import torch

    words = np.zeros((slots + 63) // 64, dtype=np.uint64)
    for s in allowed_slots:
        words[s >> 6] |= np.uint64(1) << np.uint64(s & 63)
    return torch.from_numpy(words.view(np.int64)).cuda(), len(allowed_slots)


@pytest.mark.parametrize("name", ["l2_fs_refine", "ip_fs_custom", "cos_fs_refine", "l2_sq8"])
def test_gpu_noalloc_search_matches_async_search(name):
    import torch

    lib = _lib()
    assert lib.qann_hip_noalloc_scratch_bytes() == 64 << 20
    path, q, exp, params = synth_case(name)
    idx = qann_amd.load(str(path))
    ivf = idx.base if isinstance(idx, qann_amd.RefineIndex) else idx
    kf = idx.k_factor if isinstance(idx, qann_amd.RefineIndex) else 0
    handle = idx.device_handle()
    stream = torch.cuda.current_stream().cuda_stream
    reps = 600 // q.shape[0] + 1
    qq = np.ascontiguousarray(np.tile(q, (reps, 1)))
    # the index is not cosine-normalised here: the golden queries of a cosine
    # index are prepared as the host does before either search sees them
    if idx.metric == qann_amd.Metric.Cosine:
        qq = qq / np.linalg.norm(qq.astype(np.float64), axis=1, keepdims=True).astype(np.float32)
    dq = torch.from_numpy(np.ascontiguousarray(qq, dtype=np.float32)).cuda()
    # slots of the index (ids are slots without custom ids): the filter's words cover them all
    slots = params["n"] if not params.get("custom") else None
    for k, nprobe in ((10, ivf.nprobe), (50, 7)):
        p = capi.SearchParams(k=k, nprobe=nprobe, k_factor=kf, filter_probe_scaling=1)
        for nq in (len(qq), 37):  # two calls in a row, the second smaller
            want = _run(lib, lib.qann_hip_search_async, handle, dq, nq, p, k, stream)
            got = _run(lib, lib.qann_hip_search_noalloc, handle, dq, nq, p, k, stream)
            np.testing.assert_array_equal(got[0], want[0], err_msg=f"{name} k {k} nq {nq}")
            np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
            assert (got[0][:, 0] >= 0).all()
    if slots is not None:
        rng = np.random.default_rng(5)
        live = np.asarray(idx.ids())
        for size in (len(live) // 3, 40):
            words, count = _filter_words(slots, np.sort(rng.choice(live, size, replace=False)))
            p = capi.SearchParams(k=10, nprobe=ivf.nprobe, k_factor=kf, filter_probe_scaling=1,
                                  allowed_words=words.data_ptr(), allowed_count=count)
            for nq in (len(qq), 11):
                want = _run(lib, lib.qann_hip_search_async, handle, dq, nq, p, 10, stream)
                got = _run(lib, lib.qann_hip_search_noalloc, handle, dq, nq, p, 10, stream)
                np.testing.assert_array_equal(got[0], want[0], err_msg=f"{name} filtered {size} nq {nq}")
                np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


def test_gpu_noalloc_search_chunks_large_batches_and_refuses_what_does_not_fit():
    import torch

    lib = _lib()
    path, q, exp, params = synth_case("l2_fs_refine")
    idx = qann_amd.load(str(path))
    handle = idx.device_handle()
    stream = torch.cuda.current_stream().cuda_stream
    # 20,000 queries: the coarse scores and probe arrays fit, the sums go in chunks
    qq = np.ascontiguousarray(np.tile(q, (20000 // q.shape[0] + 1, 1))[:20000])
    dq = torch.from_numpy(qq).cuda()
    p = capi.SearchParams(k=10, nprobe=idx.base.nprobe, k_factor=idx.k_factor, filter_probe_scaling=1)
    want = _run(lib, lib.qann_hip_search_async, handle, dq, len(qq), p, 10, stream)
    got = _run(lib, lib.qann_hip_search_noalloc, handle, dq, len(qq), p, 10, stream)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
    # a batch whose per-query arrays alone exceed the scratch is refused
    big = torch.zeros((400_000, qq.shape[1]), dtype=torch.float32, device="cuda")
    ids = torch.empty((400_000, 10), dtype=torch.int64, device="cuda")
    d = torch.empty((400_000, 10), dtype=torch.float32, device="cuda")
    rc = lib.qann_hip_search_noalloc(C.c_void_p(handle), C.c_void_p(big.data_ptr()), C.c_uint64(400_000),
                                     C.byref(p), C.c_void_p(ids.data_ptr()), C.c_void_p(d.data_ptr()),
                                     C.c_void_p(stream))
    assert rc == capi.EINVAL
    assert b"allocation-free" in lib.qann_hip_last_error()
    # a FlatIndex handle is refused
    fpath, fq, _, _ = synth_case("flat_ip")
    flat = qann_amd.load(str(fpath))
    fd = torch.from_numpy(np.ascontiguousarray(fq)).cuda()
    rc = lib.qann_hip_search_noalloc(C.c_void_p(flat.device_handle()), C.c_void_p(fd.data_ptr()), C.c_uint64(4),
                                     C.byref(capi.SearchParams(k=5)), C.c_void_p(ids.data_ptr()),
                                     C.c_void_p(d.data_ptr()), C.c_void_p(stream))
    assert rc == capi.EINVAL
    torch.cuda.synchronize()
