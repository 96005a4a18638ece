"""The PCA-balanced rotation that IVFIndex.train learns for enable_pq(...,
rotate=True) (detail::balanced_pca_rotation, src/rotation.cpp:71-130): the
rotation qann_amd learns on the host against the one the reference library
saves after training on the same rows, bit for bit. Host only (no GPU work).
"""
import numpy as np
import pytest

import qann_amd


@pytest.mark.parametrize("d,m,n", [(32, 16, 3000), (64, 32, 4000), (96, 12, 5000), (128, 16, 8000), (20, 4, 150)])
def test_rotation_is_the_reference_rotation(tmp_path, reference, d, m, n):
    rng = np.random.default_rng(d * 3 + m)
    c = rng.random((30, d)).astype(np.float32)
    x = (c[rng.integers(0, 30, n)] + 0.1 * rng.standard_normal((n, d))).astype(np.float32)
    idx = reference.build_ivf(x, 100, 8, metric=0, m=m, rotate=True, seed=3)
    path = tmp_path / "rot.qann"
    reference.save(idx, path)
    want = qann_amd.load(str(path)).rotation()
    got = qann_amd._qann._balanced_pca_rotation(x, m)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_rotation_is_orthonormal_and_diagonalizes_the_covariance():
    # what the mathematics fixes: R^T R = I, R^T C R diagonal, and its diagonal
    # the covariance's eigenvalues (numpy's eigvalsh)
    rng = np.random.default_rng(7)
    d, m = 48, 6
    scales = np.exp(rng.uniform(-2, 2, d))
    mix = np.linalg.qr(rng.standard_normal((d, d)))[0]
    x = ((rng.standard_normal((4000, d)) * scales) @ mix).astype(np.float32)
    r = qann_amd._qann._balanced_pca_rotation(x, m).astype(np.float64)
    np.testing.assert_allclose(r.T @ r, np.eye(d), atol=1e-5)
    cov = np.cov(x.astype(np.float64), rowvar=False)
    rot = r.T @ cov @ r
    off = rot - np.diag(np.diag(rot))
    assert np.abs(off).max() < 1e-5 * np.abs(rot).max()
    np.testing.assert_allclose(np.sort(np.diag(rot)), np.linalg.eigvalsh(cov), rtol=1e-5, atol=1e-9)
