"""GPU noise map (pdamd_estimate_noise) against the reference's outputs and
the oracle, including the 80x80x48x300 volume size."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gpu_matches_reference_fixture(golden_dir):
    from patch_denoise.space_time.utils import estimate_noise

    z = np.load(os.path.join(golden_dir, "noise_map.npz"))
    names = sorted({k[:-3] for k in z.files if k.endswith("_in")})
    for k in names:
        x, b, ref = z[k + "_in"], int(z[k + "_block"]), z[k + "_out"]
        print(f"[noise] {k} shape {x.shape} block {b}", flush=True)
        out = estimate_noise(x, b)
        np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
        tol = 1e-12 if x.dtype == np.float64 else 1e-5  # np.std accumulates in float32
        np.testing.assert_allclose(out[~np.isnan(ref)], ref[~np.isnan(ref)], rtol=tol)


def test_gpu_size_error():
    from patch_denoise.space_time.utils import estimate_noise

    with pytest.raises(IndexError):
        estimate_noise(np.zeros((4, 4, 3, 5), np.float32), 3)


def test_gpu_noise_map_80x80x48x300():
    from oracle import cpu_ref as O
    from patch_denoise.space_time.utils import estimate_noise

    rng = np.random.default_rng(8)
    x = (800.0 + 30.0 * rng.standard_normal((80, 80, 48, 300), dtype=np.float32)).astype(np.float32)
    print("[noise] 80x80x48x300 block 3", flush=True)
    out = estimate_noise(x, 3)
    assert np.isnan(out[0]).all() and np.isnan(out[:, :, -1]).all()
    inner = out[1:-1, 1:-1, 1:-1]
    assert np.isfinite(inner).all()
    assert abs(float(inner.mean()) - 30.0) < 0.05  # 2700 samples per window
    lo = (37, 40, 20)
    sub = x[lo[0]:lo[0] + 6, lo[1]:lo[1] + 6, lo[2]:lo[2] + 6]
    ref = O.estimate_noise(sub, 3)[1:-1, 1:-1, 1:-1]
    got = out[lo[0] + 1:lo[0] + 5, lo[1] + 1:lo[1] + 5, lo[2] + 1:lo[2] + 5]
    np.testing.assert_allclose(got, ref, rtol=1e-5)


def test_nordic_threshold_large_draws_on_gpu():
    """NORDIC Monte-Carlo threshold with draws above the host-SVD size (GPU
    Gram path) against the oracle's svd of the same seeded draws."""
    from oracle import cpu_ref as O
    from patch_denoise.space_time.lowrank import nordic_threshold

    np.random.seed(5)
    a = nordic_threshold(125 * 200, 200, 3)  # 5e6 elements per draw
    np.random.seed(5)
    b = O.nordic_threshold((125 * 200,), 200, 1.0, 3)
    print(f"[noise] nordic tau gpu-gram {a!r} oracle {b!r}", flush=True)
    assert a == pytest.approx(b, rel=1e-13)
