"""Top singular value of float64 matrices on the GPU (``pdamd_top_singular_values``:
FP64 Gram, Householder tridiagonalisation, bisection), the kernel behind the
NORDIC and QUT Monte-Carlo thresholds (reference: space_time/lowrank.py:221-232,
:446-456, ``max(svd(draw, compute_uv=False))``).

* Matrices with prescribed singular values (U diag(s) V^T from QR factors):
  the top value to 1e-13 relative, tall, wide and square, Gram sizes 2 .. 600
  (the one-wave, four-wave and large-Gram Householder kernels).
* Gaussian draws against ``numpy.linalg.svd``: 1e-13 relative.
* Repeat calls give identical bits (the split contraction is summed in a
  fixed order).
* The QUT draw helper against the reference's per-draw SVD loop.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _prescribed(rng, rows, cols, top):
    n = min(rows, cols)
    u, _ = np.linalg.qr(rng.standard_normal((rows, n)))
    v, _ = np.linalg.qr(rng.standard_normal((cols, n)))
    s = np.sort(rng.uniform(0.1, 0.9, n))[::-1] * top
    s[0] = top
    return (u * s) @ v.T


@pytest.mark.parametrize("rows,cols,batch", [
    (5, 2, 3), (2, 7, 2), (40, 40, 4), (300, 64, 5), (64, 300, 5), (125, 300, 7),
    (1000, 200, 2), (700, 520, 1), (600, 900, 1)])
def test_prescribed_top_singular_value(rows, cols, batch):
    from patch_denoise import _native as N

    rng = np.random.default_rng(rows * 7 + cols)
    tops = rng.uniform(1.0, 50.0, batch)
    a = np.stack([_prescribed(rng, rows, cols, t) for t in tops])
    sv = N.top_singular_values(a)
    err = np.max(np.abs(sv - tops) / tops)
    print(f"[topsv] prescribed {batch} x {rows}x{cols}: max rel err {err:.2e}", flush=True)
    assert err < 1e-13


def test_gaussian_draws_match_numpy_svd_and_repeat():
    from patch_denoise import _native as N

    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, 3000, 150))
    sv1 = N.top_singular_values(a)
    sv2 = N.top_singular_values(a)
    ref = np.array([np.linalg.svd(x, compute_uv=False)[0] for x in a])
    err = np.max(np.abs(sv1 - ref) / ref)
    print(f"[topsv] gaussian 6 x 3000x150: max rel err vs numpy svd {err:.2e}", flush=True)
    assert err < 1e-13
    assert np.array_equal(sv1, sv2)


def test_qut_draws_match_reference_loop():
    from patch_denoise.space_time.lowrank import _qut_max_svals
    from scipy.linalg import svd

    np.random.seed(9)
    got = _qut_max_svals(27, 40, 2.5, 50)
    np.random.seed(9)
    ref = np.array([np.max(svd(np.random.randn(27, 40) * 2.5, compute_uv=False))
                    for _ in range(50)])
    assert got.shape == (50,)
    assert np.max(np.abs(got - ref) / ref) < 1e-13


def test_top_singular_values_rejects_bad_shapes():
    from patch_denoise import _native as N

    with pytest.raises(N.NativeError):
        N.top_singular_values(np.ones((1, 50)))
    with pytest.raises(N.NativeError):
        N.top_singular_values(np.ones((2, 2049, 2049))[:, :2049, :2049])
