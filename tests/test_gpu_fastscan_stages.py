"""The fast-scan's two integer stages on the GPU, checked directly against the
oracle's quantize_lut and fast_scan_block_mask (themselves pinned to the
reference's known answers in tests/test_oracle.py): LUT bytes, scale and
offset bit for bit, and every u16 sum, by the device code the search kernels
run (qann_hip_fs_quantize_lut / qann_hip_fs_block_sums)."""
import ctypes as C

import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu

SUBSPACES = [1, 2, 3, 4, 5, 8, 16, 24, 32, 48, 50, 63, 64, 65, 96, 128, 129, 192, 384, 392, 1023, 1024]


def tables(rng, m):
    """Tables of several kinds: uniform, tiny and huge ranges, integer entries
    whose scaled offsets land on halves (ties), constant rows, zero rows with
    both signs of zero, negative rows."""
    out = [rng.uniform(-5, 5, (m, 16)), rng.uniform(0, 1e-3, (m, 16)), rng.uniform(-1e4, 1e4, (m, 16))]
    t = rng.integers(0, 127, (m, 16)).astype(np.float64)
    t[:, 0], t[0, 1] = 0.0, 126.0  # widest 126: scale 0.5, odd entries round half to even
    out.append(t)
    c = np.repeat(rng.uniform(-3, 3, (m, 1)), 16, axis=1)
    out.append(c)  # constant rows: scale 1, offset the sum of the rows
    z = rng.uniform(0, 2, (m, 16))
    z[:, 3] = 0.0
    z[::2, 7] = -0.0
    out.append(z)
    out.append(-rng.uniform(10, 160, (m, 16)))
    return np.ascontiguousarray(np.stack(out).astype(np.float32))


@pytest.mark.parametrize("m", SUBSPACES)
def test_gpu_quantize_lut_matches_oracle(oracle, m):
    rng = np.random.default_rng(m)
    t = tables(rng, m)
    n = t.shape[0]
    lib = capi.lib()
    lut = np.zeros((n, m * 16), np.uint8)
    scale = np.zeros(n, np.float32)
    offset = np.zeros(n, np.float32)
    rc = lib.qann_hip_fs_quantize_lut(capi.ptr(t), C.c_uint64(n), C.c_uint32(m), capi.ptr(lut), capi.ptr(scale),
                                      capi.ptr(offset))
    assert rc == 0, lib.qann_hip_last_error()
    for i in range(n):
        want_lut, want_s, want_o = oracle.quantize_lut(t[i], m)
        np.testing.assert_array_equal(lut[i], want_lut, err_msg=f"m {m} table {i}")
        assert scale[i].view(np.uint32) == np.float32(want_s).view(np.uint32), (m, i, scale[i], want_s)
        assert offset[i].view(np.uint32) == np.float32(want_o).view(np.uint32), (m, i, offset[i], want_o)


@pytest.mark.parametrize("m", SUBSPACES)
def test_gpu_block_sums_match_oracle(oracle, m):
    rng = np.random.default_rng(1000 + m)
    nb = 24
    luts = rng.integers(0, 64, (nb, m * 16), dtype=np.uint8)
    luts[0] = 63  # the largest sums: m * 63
    blocks = rng.integers(0, 256, (nb, m * 16), dtype=np.uint8)
    blocks[1] = 0
    lib = capi.lib()
    sums = np.zeros((nb, 32), np.uint16)
    rc = lib.qann_hip_fs_block_sums(capi.ptr(luts), C.c_uint32(m), capi.ptr(blocks), C.c_uint64(nb), capi.ptr(sums))
    assert rc == 0, lib.qann_hip_last_error()
    for b in range(nb):
        _, want = oracle.fast_scan_block_mask(luts[b], blocks[b], m, 0)
        np.testing.assert_array_equal(sums[b], want, err_msg=f"m {m} block {b}")
