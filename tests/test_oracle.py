"""Pins the CPU oracle (oracle/qann_oracle.cpp) before anything is checked
against it:

* the reference's own saved indexes and the results it recorded for them
  (tests/data of maxtmiller/qann: ids exact, distances within 1e-5 as its
  tests/test_python.py:1108-1118, 1776-1810 demand);
* the synthetic golden fixtures written by the reference library itself
  (tests/golden/make_golden.py): ids and distance bits exact;
* the reference library loaded here (oracle/_ref), where it is built;
* the known answers of the reference's tests/test_fast_scan.cpp and closed
  forms of the arithmetic (integer sums, half-to-even rounding, error bound).
"""
import numpy as np
import pytest

from cases import REF_FIXTURES, REF_FIXTURES_OTHER, SYNTH_CASES, ref_fixture, synth_case, synth_queries


@pytest.mark.parametrize("name,queries", REF_FIXTURES + REF_FIXTURES_OTHER)
def test_oracle_reproduces_reference_fixtures(oracle, name, queries):
    path, q, ids, dists = ref_fixture(name, queries)
    idx = oracle.load(path)
    got_ids, got_d = idx.batch_query(q, ids.shape[1])
    np.testing.assert_array_equal(got_ids, ids)
    np.testing.assert_allclose(got_d, dists, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name", SYNTH_CASES)
def test_oracle_bit_exact_on_reference_goldens(oracle, name):
    path, q, exp, params = synth_case(name)
    idx = oracle.load(path)
    for npb, k, fname, allowed, stem in synth_queries(exp, params):
        if npb is not None:
            idx.set_nprobe(npb)
        ids, d = idx.batch_query(q, k, allowed=allowed)
        np.testing.assert_array_equal(ids, exp[stem + "_ids"], err_msg=stem)
        np.testing.assert_array_equal(d.view(np.uint32), exp[stem + "_dists"].view(np.uint32), err_msg=stem)


@pytest.mark.parametrize("name,queries", REF_FIXTURES + REF_FIXTURES_OTHER)
def test_oracle_bit_exact_with_reference_library(oracle, reference, name, queries):
    path, q, _, _ = ref_fixture(name, queries)
    a = oracle.load(path).batch_query(q, 10)
    b = reference.load(path).batch_query(q, 10)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@pytest.mark.parametrize("n", list(range(1, 70)) + [96, 100, 127, 128, 129, 768, 1000])
def test_exact_distance_order_matches_reference(oracle, reference, n):
    # the AVX2 accumulator order and g++'s tail (4 unfused, then fused) of
    # l2_distance / inner_product, bit for bit
    rng = np.random.default_rng(n)
    for _ in range(4):
        a = rng.standard_normal(n).astype(np.float32)
        b = rng.standard_normal(n).astype(np.float32)
        assert oracle.l2_distance(a, b) == reference.l2_distance(a, b)
        assert oracle.inner_product(a, b) == reference.inner_product(a, b)


def test_exact_distance_against_float64(oracle):
    # the restated order is a float32 sum of the true terms: within a few ulps
    # of the float64 value for positive terms
    rng = np.random.default_rng(5)
    for n in (7, 32, 100, 768):
        a = rng.standard_normal(n).astype(np.float32)
        b = rng.standard_normal(n).astype(np.float32)
        exact = float(np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2))
        assert abs(float(oracle.l2_distance(a, b)) - exact) <= 1e-6 * exact * n


# --- quantize_lut: tests/test_fast_scan.cpp known answers ------------------


def test_quantize_lut_constant_rows(oracle):
    lut, scale, offset = oracle.quantize_lut(np.full(4 * 16, 3.0, np.float32), 4)
    assert scale == np.float32(1.0) and offset == np.float32(12.0)
    assert (lut == 0).all()


def test_quantize_lut_negative_rows(oracle):
    t = np.zeros(32, np.float32)
    t[:16] = -500.0 + 10.0 * np.arange(16)
    t[16:] = -40.0 - 2.0 * np.arange(16)
    lut, scale, offset = oracle.quantize_lut(t, 2)
    assert scale == np.float32(63) / np.float32(150)
    assert offset == np.float32(-570)
    assert lut[0] == 0 and lut[15] == 63 and lut[31] == 0
    assert lut[16] == np.rint(np.float32(30) * scale)


def test_quantize_lut_rounds_half_to_even(oracle):
    t = np.zeros(32, np.float32)
    t[1] = 63.0
    probes = [0.5, 1.5, 2.5, 62.5, np.nextafter(np.float32(0.5), np.float32(0)),
              np.nextafter(np.float32(0.5), np.float32(1)), np.nextafter(np.float32(1.5), np.float32(0)), 0.25, 3.0]
    t[17:26] = np.array(probes, np.float32)
    lut, scale, _ = oracle.quantize_lut(t, 2)
    assert scale == np.float32(1.0)
    assert lut[17:26].tolist() == [0, 2, 2, 62, 0, 1, 1, 0, 3]


def test_quantize_lut_error_bound(oracle):
    m = 32
    rng = np.random.default_rng(7)
    t = rng.uniform(0, 5000, m * 16).astype(np.float32)
    for s in range(m):
        t[s * 16:(s + 1) * 16] = t[s * 16:(s + 1) * 16] * (1 + s % 3) + 100.0 * s
    lut, scale, offset = oracle.quantize_lut(t, m)
    rows = lut.reshape(m, 16)
    assert (rows.min(axis=1) == 0).all() and lut.max() == 63
    bound = m / (2 * scale) + 1e-2
    for _ in range(200):
        j = rng.integers(0, 16, m)
        exact = float(t.reshape(m, 16)[np.arange(m), j].astype(np.float64).sum())
        approx = rows[np.arange(m), j].sum() / scale + offset
        assert abs(approx - exact) <= bound


@pytest.mark.parametrize("m", [2, 6, 16, 64, 196, 1024])
def test_quantize_lut_matches_reference(oracle, reference, m):
    rng = np.random.default_rng(m + 100)
    for t in (rng.uniform(-3000, 3000, m * 16).astype(np.float32), np.full(m * 16, -2.5, np.float32)):
        a = oracle.quantize_lut(t, m)
        b = reference.quantize_lut(t, m)
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1] == b[1] and a[2] == b[2]


# --- fast-scan block sums and masks ----------------------------------------


def fs_block(codes):
    """Packs (32, m) 4-bit codes into one block of the reference's layout (fs_set)."""
    m = codes.shape[1]
    block = np.zeros(m * 16, np.uint8)
    for v in range(32):
        for s in range(m):
            b = s * 16 + v % 16
            block[b] = (block[b] & 0xF0) | codes[v, s] if v < 16 else (block[b] & 0x0F) | (codes[v, s] << 4)
    return block


def test_block_layout_known_bytes():
    codes = np.zeros((32, 2), np.uint8)
    codes[0, 0], codes[16, 0], codes[5, 1], codes[31, 1] = 0x3, 0xA, 0x7, 0x1
    block = fs_block(codes)
    assert block[0] == 0xA3 and block[16 + 5] == 0x07 and block[16 + 15] == 0x10


@pytest.mark.parametrize("m", [2, 4, 6, 16, 64, 196, 1024])
def test_fast_scan_sums_closed_form(oracle, m):
    rng = np.random.default_rng(m)
    codes = rng.integers(0, 16, (32, m), dtype=np.uint8)
    for lut in (rng.integers(0, 64, m * 16, dtype=np.uint8), np.full(m * 16, 63, np.uint8)):
        mask, sums = oracle.fast_scan_block_mask(lut, fs_block(codes), m, 65535)
        expected = lut.reshape(m, 16)[np.arange(m)[None, :], codes].astype(np.int64).sum(axis=1)
        np.testing.assert_array_equal(sums, expected)
        assert expected.max() <= 64512
        srt = np.sort(expected)
        for limit in (0, 65535, int(srt[0]), int(srt[10]), int(srt[31]), int(srt[20]) + 1):
            mask, _ = oracle.fast_scan_block_mask(lut, fs_block(codes), m, limit)
            want = sum(1 << v for v in range(32) if expected[v] < limit)
            assert mask == want
