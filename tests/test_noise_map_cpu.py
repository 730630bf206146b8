"""CPU checks of the noise-map oracle (space_time/utils.py:207-242 restated):
against the reference's own outputs (tests/golden/noise_map.npz) and against
an independent float64 box-sum computation."""

import os

import numpy as np
import pytest


def _cases(golden_dir):
    z = np.load(os.path.join(golden_dir, "noise_map.npz"))
    names = sorted({k[:-3] for k in z.files if k.endswith("_in")})
    return z, names


def test_oracle_matches_reference_fixture(golden_dir):
    from oracle import cpu_ref as O

    z, names = _cases(golden_dir)
    assert len(names) == 6
    for k in names:
        out = O.estimate_noise(z[k + "_in"], int(z[k + "_block"]))
        ref = z[k + "_out"]
        np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
        np.testing.assert_array_equal(out[~np.isnan(ref)], ref[~np.isnan(ref)])


def _boxsum_std(x, b):
    """Independent: per-voxel temporal moments, then b^d box sums by
    cumulative sums (float64); returns the centre-indexed map (NaN elsewhere)."""
    sp = x.shape[:-1]
    T = x.shape[-1]
    s1 = x.sum(-1).astype(np.complex128)
    s2 = (np.abs(x.astype(np.complex128)) ** 2).sum(-1)

    def box(a):
        for ax in range(a.ndim):
            c = np.cumsum(a, axis=ax)
            pad = [(0, 0)] * a.ndim
            pad[ax] = (1, 0)
            c = np.pad(c, pad)
            sl_hi = [slice(None)] * a.ndim
            sl_lo = [slice(None)] * a.ndim
            sl_hi[ax] = slice(b, None)
            sl_lo[ax] = slice(0, -b)
            a = c[tuple(sl_hi)] - c[tuple(sl_lo)]
        return a

    N = b ** len(sp) * T
    m1, m2 = box(s1) / N, box(s2) / N
    std = np.sqrt(np.maximum(m2 - np.abs(m1) ** 2, 0))
    out = np.full(sp, np.nan)
    out[tuple(slice(b // 2, s - b + b // 2 + 1) for s in sp)] = std
    return out


def test_oracle_matches_box_sums(golden_dir):
    from oracle import cpu_ref as O

    z, names = _cases(golden_dir)
    for k in names:
        x, b = z[k + "_in"], int(z[k + "_block"])
        tol = 1e-12 if x.dtype == np.float64 else 1e-5  # np.std accumulates in float32
        np.testing.assert_allclose(O.estimate_noise(x, b), _boxsum_std(x, b), rtol=tol)


def test_oracle_size_errors():
    from oracle import cpu_ref as O

    with pytest.raises(IndexError):
        O.estimate_noise(np.zeros((3, 5, 4, 6), np.float32), 3)  # 3 <= block
