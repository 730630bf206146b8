"""Gather recombination (every output element sums its covering patches in
increasing patch index, no atomics) against the atomic scatter and the oracle.

* The 80x80x48x300 MP-PCA volume: two gather runs are bit-identical; gather
  and scatter agree to float32 rounding (rel-L2 <= 1e-6) with identical ranks.
* The 32x32x16x50 workload with 3^3 patches (no mask, 11,760 patches; M-side:
  27 voxels < 50 frames): the float32 weights equal
  the oracle's bit for bit -- they are summed in the reference's order with
  its rounding (base.py:243-256: f32(f64(w) + theta) per patch) -- and the
  output matches it to 1e-5.
* Hybrid (a Python-float theta: float32 additions) with a mask: weights equal
  the oracle's bit for bit.
* T-side patches (fewer frames than voxels: left factors Xc V diag(wc), right
  factors theta [m; V^T]): two gather runs are bit-identical and agree with
  the scatter to float32 rounding with identical ranks (7^3 x 200, Gram size
  200; 5^3 x 60, Gram size 60); on the 32x32x16x50 workload with 5^3 patches
  the weights equal the oracle's bit for bit and the output matches it to
  1e-5.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.nan_to_num(np.asarray(a, np.float64)), np.nan_to_num(np.asarray(b, np.float64))
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _run(den, data, mask, mode):
    from patch_denoise import _native as N

    prob, p_s, p_o, mask_arr = den._prepare(data, mask, 50)
    var_map = den._configure(prob, data, p_s)
    prob.recombine_mode = N.RECOMBINE[mode]
    run = N.DeviceRun(prob, data, mask_arr, var_map, want_patch_info=True).run()
    out, w = run.result(data.shape)
    return out, w, run.patch_info()


def test_gather_repeats_bitwise_and_matches_scatter():
    from patch_denoise import MPPCADenoiser, workloads

    data, mask = workloads.config2()
    den = MPPCADenoiser(5, 4, threshold_scale=1.0)
    print("[gather] 80x80x48x300 gather x2, scatter x1", flush=True)
    g1, w1, i1 = _run(den, data, mask, "gather")
    g2, w2, _ = _run(den, data, mask, "gather")
    s1, ws, is_ = _run(den, data, mask, "scatter")
    assert np.array_equal(g1, g2, equal_nan=True) and np.array_equal(w1, w2)
    assert np.array_equal(i1["rank"], is_["rank"])
    e, ew = rel_l2(g1, s1), rel_l2(w1, ws)
    print(f"[gather] gather vs scatter: out {e:.2e} weights {ew:.2e}", flush=True)
    assert e <= 1e-6 and ew <= 1e-6


def test_gather_weights_bit_exact_against_oracle():
    from oracle import cpu_ref as O
    from patch_denoise import MPPCADenoiser, workloads

    data = workloads.config1()
    out, w, info = _run(MPPCADenoiser(3, 2, threshold_scale=1.0), data, None, "gather")
    ref, ref_w, _, _ = O.mp_pca(data, 3, 2)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    nbad = int(np.sum(w != ref_w))
    e = rel_l2(out, ref)
    print(f"[gather] 32x32x16x50: weights differing {nbad}, out rel-L2 {e:.2e}", flush=True)
    assert nbad == 0 and e <= 1e-5


def test_gather_hybrid_masked_weights_bit_exact():
    from oracle import cpu_ref as O
    from patch_denoise import HybridPCADenoiser, workloads
    from patch_denoise.space_time.base import spatial_variance_map

    rng = np.random.default_rng(5)
    shape, T = (18, 16, 14), 40
    sig = 200 + 15 * np.tensordot(rng.standard_normal((3,) + shape), rng.standard_normal((3, T)),
                                  axes=([0], [0]))
    data = (sig + 6 * rng.standard_normal(shape + (T,))).astype(np.float32)
    mask = workloads._ellipsoid(shape, (8, 7, 6))
    sigma = 6.0 * (1 + 0.2 * rng.random(shape))
    den = HybridPCADenoiser(3, 2)
    den._var_map = spatial_variance_map(sigma, data.shape)
    out, w, _ = _run(den, data, mask, "gather")
    ref, ref_w, _, _ = O.hybrid_pca(data, 3, 2, mask=mask, noise_std=sigma)
    nbad = int(np.sum(w != ref_w))
    e = rel_l2(out, ref)
    print(f"[gather] hybrid masked: weights differing {nbad}, out rel-L2 {e:.2e}", flush=True)
    assert nbad == 0 and e <= 1e-5


def _tside_volume(shape, T, seed, rank=4):
    rng = np.random.default_rng(seed)
    sig = 600 + 40 * np.tensordot(rng.standard_normal((rank,) + shape),
                                  rng.standard_normal((rank, T)), axes=([0], [0]))
    return (sig + 25 * rng.standard_normal(shape + (T,))).astype(np.float32)


@pytest.mark.parametrize("shape,T,p", [((11, 12, 10), 200, 7), ((12, 12, 10), 100, 5),
                                       ((12, 12, 10), 60, 5)],
                         ids=["p7_T200", "p5_T100", "p5_T60"])
def test_gather_tside_repeats_bitwise_and_matches_scatter(shape, T, p):
    from patch_denoise import MPPCADenoiser

    data = _tside_volume(shape, T, 7)
    den = MPPCADenoiser(p, p - 1, threshold_scale=1.0)
    print(f"[gather] T-side {shape}x{T} p{p}: gather x2, scatter x1", flush=True)
    g1, w1, i1 = _run(den, data, None, "gather")
    g2, w2, _ = _run(den, data, None, "gather")
    s1, ws, is_ = _run(den, data, None, "scatter")
    assert np.array_equal(g1, g2, equal_nan=True) and np.array_equal(w1, w2)
    assert np.array_equal(i1["rank"], is_["rank"])
    e, ew = rel_l2(g1, s1), rel_l2(w1, ws)
    print(f"[gather] T-side gather vs scatter: out {e:.2e} weights {ew:.2e}", flush=True)
    assert e <= 1e-6 and ew <= 1e-6


def test_gather_tside_weights_bit_exact_against_oracle():
    from oracle import cpu_ref as O
    from patch_denoise import MPPCADenoiser, workloads

    data = workloads.config1()
    out, w, info = _run(MPPCADenoiser(5, 2, threshold_scale=1.0), data, None, "gather")
    ref, ref_w, _, _ = O.mp_pca(data, 5, 2)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    nbad = int(np.sum(w != ref_w))
    e = rel_l2(out, ref)
    print(f"[gather] T-side 32x32x16x50 p5: weights differing {nbad}, out rel-L2 {e:.2e}",
          flush=True)
    assert nbad == 0 and e <= 1e-5


def test_gather_tside_gram100_matches_oracle():
    """Frame-side Gram of 100 (125 voxels > 100 frames, float32): the 16 x 16
    MFMA Gram tiles, the one-wavefront Householder, the eigen kernel and the
    T-side gather against the oracle; weights bit-exact, output <= 1e-5."""
    from oracle import cpu_ref as O
    from patch_denoise import MPPCADenoiser

    data = _tside_volume((12, 12, 10), 100, 11)
    out, w, info = _run(MPPCADenoiser(5, 4, threshold_scale=1.0), data, None, "gather")
    ref, ref_w, _, _ = O.mp_pca(data, 5, 4)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    nbad = int(np.sum(w != ref_w))
    e = rel_l2(out, ref)
    print(f"[gather] T-side 12x12x10x100 p5: weights differing {nbad}, out rel-L2 {e:.2e}",
          flush=True)
    assert nbad == 0 and e <= 1e-5


def test_gather_refused_where_ineligible():
    """Forcing the gather on a rule other than MP-PCA / Hybrid (raw SVT) is an error."""
    from patch_denoise import RawSVDDenoiser, _native as N

    data = np.random.default_rng(1).standard_normal((12, 12, 10, 60)).astype(np.float32) + 50
    den = RawSVDDenoiser(5, 4, threshold_value=3.0)
    den._threshold = 3.0  # set by denoise() from threshold_value * threshold_scale
    prob, p_s, p_o, mask_arr = den._prepare(data, None, 50)
    den._configure(prob, data, p_s)
    prob.recombine_mode = N.RECOMBINE["gather"]
    with pytest.raises(N.NativeError, match="gather recombination"):
        N.DeviceRun(prob, data, None, None)
