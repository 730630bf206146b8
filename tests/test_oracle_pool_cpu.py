"""The parallel oracle helper (oracle/pool.py) against the oracle's own
sequential patch driver on small volumes: same ranks, same window output."""

import numpy as np

from oracle import pool as OP


def test_pool_matches_sequential_oracle():
    from oracle import cpu_ref as O
    from patch_denoise.workloads import _ellipsoid

    rng = np.random.default_rng(7)
    shape, T = (9, 8, 7), 30
    sig = 100 + 8 * np.tensordot(rng.standard_normal((2,) + shape), rng.standard_normal((2, T)),
                                 axes=([0], [0]))
    data = (sig + 3 * rng.standard_normal(shape + (T,))).astype(np.float32)
    mask = _ellipsoid(shape, (4, 4, 3))
    var = (3.0 * (1 + 0.1 * rng.random(shape))) ** 2
    ps, po = (3, 3, 3, T), (2, 2, 2, T)
    sel, _ = O.select_patches(mask, data.shape, ps, po, 50)
    grid = O.StridedPatches(data, ps, po).grid
    cases = [
        ("mppca", {"ts": 1.0}, O.mp_pca(data, 3, 2, mask=mask)),
        ("hybrid", {"var": var}, O.hybrid_pca(data, 3, 2, mask=mask, noise_std=np.sqrt(var))),
        ("raw", {"tau": 40.0}, O.raw_svt(data, 3, 2, mask=mask, threshold=40.0)),
    ]
    lo, hi = (2, 2, 2), (6, 6, 5)
    pool = OP._pool()
    try:
        with OP.SharedVolume(data) as sv:
            for method, params, (ref, ref_w, _, _) in cases:
                ranks, margin = OP.oracle_ranks(sv, ps, po, method, params, sel, pool)
                assert np.all(np.isfinite(margin)) and np.all(margin >= 0)
                out, w, wr = OP.oracle_window(sv, ps, po, method, params, sel, grid, lo, hi, pool)
                assert all(ranks[np.searchsorted(sel, L)] == r for L, r in wr.items())
                win = ref[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
                m = mask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
                out = np.where(m[..., None], out, 0.0)
                assert np.array_equal(np.isnan(out), np.isnan(win))
                np.testing.assert_allclose(np.nan_to_num(out), np.nan_to_num(win), rtol=2e-5,
                                           atol=1e-3)
                np.testing.assert_allclose(w, ref_w[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]],
                                           rtol=1e-6)
                # the reference's float32 weight sums, bit for bit
                w32 = OP.window_weights32(ps, po, grid, wr, lo, hi, T)
                assert np.array_equal(w32, ref_w[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]])
    finally:
        pool.shutdown()


def test_mppca_margin_flags_a_planted_tie():
    """A spectrum whose top eigenvalue is solved for so that the rule's first
    comparison (lowrank.py:48-49: mean * 4 sqrt(11 / M) against a[9] - a[0])
    is an exact tie has a margin at float32 rounding; the same spectrum with
    the top value moved by 1% has a margin of about that size."""
    M, T = 125, 300
    bulk = np.array([0.98, 0.99, 1.0, 1.0, 1.01, 1.02, 1.03, 1.04, 1.05])
    c = 4 * np.sqrt(11 / M)
    x = (c * bulk.sum() / 10 + bulk[0]) / (1 - c / 10)  # mean(a) c == x - a[0]
    tie = np.append(bulk, x).astype(np.float32)
    assert OP._mppca_margin(tie, M, T, 1.0) < 1e-6
    far = np.append(bulk, 1.01 * x).astype(np.float32)
    assert 1e-3 < OP._mppca_margin(far, M, T, 1.0) < 2e-2
