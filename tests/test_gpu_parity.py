"""GPU parity: the HIP path (through the C ABI) against the reference fixtures
and the CPU oracle.

Tolerances: outputs are compared by relative L2 over finite entries with the
NaN pattern required to match exactly.  float32 data: rel-L2 <= 1e-5 when
every patch keeps the reference's rank.  float64 data: 5e-7 -- the
reference divides its float64 sums by a float32 weight array accumulated in
C-order patch sequence, and our atomics add the same terms in another order,
so the divisor itself differs by float32 ulps.
"""

import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E2E = sorted(glob.glob(os.path.join(GOLDEN, "e2e_*.npz")))


def _pd():
    import patch_denoise

    return patch_denoise


def rel_l2(got, want):
    fin = np.isfinite(want)
    d = np.asarray(got, np.complex128)[fin] - np.asarray(want, np.complex128)[fin]
    den = np.linalg.norm(np.asarray(want, np.complex128)[fin])
    return float(np.linalg.norm(d) / max(den, 1e-300))


def _spec(z):
    return {k: z[k] for k in ("sigma", "sigma_map", "raw_tau", "optimal_ok")}


@pytest.mark.parametrize("path", E2E, ids=[os.path.basename(p)[4:-4] for p in E2E])
def test_e2e_against_reference_fixtures(path):
    from .golden.cases import run_table

    pd = _pd()
    z = np.load(path)
    data = z["input"]
    mask = z["mask"] if "mask" in z.files else None
    ps, po = z["patch_shape"], z["patch_overlap"]
    ps = int(ps) if ps.ndim == 0 else tuple(int(v) for v in ps)
    po = int(po) if po.ndim == 0 else tuple(int(v) for v in po)
    common = dict(patch_shape=ps, patch_overlap=po, mask=mask,
                  mask_threshold=int(z["mask_threshold"]))
    tol = 5e-7 if data.dtype in (np.float64, np.complex128) else 1e-5
    failures = []
    for key, fn_name, kw in run_table(_spec(z)):
        if fn_name == "nordic":
            np.random.seed(0)
        print(f"[gpu-e2e] {os.path.basename(path)} {key}", flush=True)
        out, w, noise, rank = getattr(pd, fn_name)(data, **common, **kw)
        want, want_w = z[key + "__out"], z[key + "__w"]
        assert out.dtype == want.dtype and out.shape == want.shape
        if not np.array_equal(np.isnan(out), np.isnan(want)):
            failures.append(f"{key}: NaN pattern")
            continue
        e = rel_l2(out, want)
        ew = rel_l2(w, want_w)
        print(f"[gpu-e2e]   rel-L2 out {e:.3g} weights {ew:.3g}", flush=True)
        if e > tol or ew > 1e-6:
            failures.append(f"{key}: rel-L2 out {e:.3g} weights {ew:.3g}")
    assert not failures, failures


def test_per_patch_cfg1_against_fixture():
    """Casorati-batch path (plugin method): spectra, ranks, noise, patches."""
    pd = _pd()
    z = np.load(os.path.join(GOLDEN, "patches_cfg1.npz"))
    den = pd.MPPCADenoiser(5, 2, threshold_scale=1.0)
    out, ranks, noise, eig = den._process_casorati(
        z["casorati"], configure=lambda prob, p: den._set_scale(prob, p.dtype, 1.0))
    np.testing.assert_array_equal(ranks, z["mppca_rank"])
    want_eig = z["eig_top10"]
    scale = np.abs(want_eig).max(axis=1, keepdims=True)
    assert np.abs(eig - want_eig).max() <= 2e-6 * scale.max()
    np.testing.assert_allclose(noise, z["mppca_noisevar"], rtol=2e-5)
    for j in range(len(out)):
        assert rel_l2(out[j], z["mppca_patch"][j]) < 1e-5


def test_plugin_patch_processing_matches_oracle():
    """``_patch_processing`` of every rule on single Casorati matrices."""
    from oracle import cpu_ref as O

    pd = _pd()
    rng = np.random.default_rng(5)
    for M, T in ((27, 20), (36, 20), (20, 36), (125, 50)):
        base = rng.standard_normal((M, 3)) @ rng.standard_normal((3, T)) * 8
        X = (100 + base + rng.standard_normal((M, T)) * 2).astype(np.float32)
        got = pd.MPPCADenoiser(3, 2, threshold_scale=1.0)._patch_processing(X, threshold_scale=1.0)
        want = O.mppca_patch(X.copy(), 1.0)
        assert got[1] == want[1]
        assert rel_l2(got[0], want[0]) < 1e-5
        tau = float(2.0 * (np.sqrt(M) + np.sqrt(T)))
        r = pd.RawSVDDenoiser(3, 2, threshold_value=tau)
        r._threshold = tau
        got = r._patch_processing(X)
        want = O.raw_patch(X.copy(), tau)
        assert got[1] == want[1]
        assert rel_l2(got[0], want[0]) < 1e-5
        if T <= M:
            mpm = O.mp_median(T / M)
            for loss in ("fro", "nuc", "ope"):
                o = pd.OptimalSVDDenoiser(3, 2, loss=loss)
                got = o._patch_processing(X, shrink_func=loss, mp_median=mpm)
                want = O.optimal_patch(X.copy(), O.SHRINKERS[loss], mpm)
                assert got[1] == want[1], (M, T, loss)
                assert rel_l2(got[0], want[0]) < 1e-5, (M, T, loss)


def test_selection_matches_oracle():
    from oracle import cpu_ref as O
    from patch_denoise import _native as N
    from patch_denoise.space_time.base import _geometry

    rng = np.random.default_rng(3)
    shape = (11, 9, 7, 13)
    for trial in range(6):
        ps = tuple(int(v) for v in rng.integers(2, 5, 3)) + (13,)
        po = tuple(int(rng.integers(0, p)) for p in ps[:3]) + (13,)
        if trial % 2:
            mask = rng.random(shape[:-1]) > 0.35
        else:
            mask = rng.random(shape) > 0.3
        thr = float(rng.choice([-1, 30, 50, 70]))
        want, _ = O.select_patches(mask, shape, ps, po, thr)
        s4, p4, st4 = _geometry(shape, ps, po)
        prob = N.make_problem(s4, p4, st4, np.float32, N.MPPCA)
        prob.mask_kind = N.MASK_FULL if mask.ndim == 4 else N.MASK_SPATIAL
        prob.mask_threshold = thr
        got = N.select_patches(prob, mask)
        np.testing.assert_array_equal(got, want)


def test_cfg1_full_run_against_fixture():
    from patch_denoise import workloads

    pd = _pd()
    z = np.load(os.path.join(GOLDEN, "cfg1_mppca.npz"))
    data = workloads.config1()
    den = pd.MPPCADenoiser(5, 2, threshold_scale=1.0)
    out, w, _, _ = den.denoise(data, progbar=False)
    assert int(np.isnan(out).sum()) == int(z["out_nan_count"])
    flat = z["sample_index"]
    got = out.reshape(-1)[flat]
    assert rel_l2(got, z["sample_out"]) < 1e-5
    np.testing.assert_allclose(w.reshape(-1)[flat], z["sample_w"], rtol=1e-6)
    # whole-volume checksums of the reference output
    s = np.nansum(out.astype(np.float64))
    assert abs(s - float(z["out_nansum"])) <= 1e-6 * abs(float(z["out_nansum"]))


def test_split_bf16_tside_gram_against_oracle():
    """precision="bf16x3" on a T-side Gram (5^3 voxels > 100 frames, Gram 100):
    split-BF16 three-product MFMA with FP32 accumulation, hi / lo planes
    stored frame-major.  Bound 1e-3 against the oracle (the split drops the
    lo x lo term, ~2^-16 relative)."""
    from oracle import cpu_ref as O

    pd = _pd()
    rng = np.random.default_rng(12)
    shape, T = (11, 10, 9), 100
    sig = 500 + 20 * np.tensordot(rng.standard_normal((4,) + shape),
                                  rng.standard_normal((4, T)), axes=([0], [0]))
    data = (sig + 10 * rng.standard_normal(shape + (T,))).astype(np.float32)
    out, w, _, _ = pd.mp_pca(data, 5, 4, precision="bf16x3")
    ref32, _, _, _ = pd.mp_pca(data, 5, 4, precision="fp32")
    want, want_w, _, _ = O.mp_pca(data, 5, 4)
    e, e32 = rel_l2(out, want), rel_l2(out, ref32)
    print(f"[gpu-bf16] T-side Gram 100: rel-L2 vs oracle {e:.3g}, vs fp32 path {e32:.3g}", flush=True)
    assert np.array_equal(np.isnan(out), np.isnan(want))
    assert e <= 1e-3 and e32 > 0.0  # the split path really ran (not bit-identical to FP32)


def test_float32_noise_map_against_reference_fixture():
    """Hybrid-PCA and optimal shrinkage with a float32 noise-level map: the
    reference averages each patch's prior in float32 (the map's dtype); the
    device rounds its float64 patch mean to float32 the same way."""
    pd = _pd()
    z = np.load(os.path.join(GOLDEN, "map32_prior.npz"))
    for name in ("tside_mask", "2d_p6"):
        data, sm32 = z[name + "__input"], z[name + "__sigma_map32"]
        assert sm32.dtype == np.float32
        mask = z[name + "__mask"] if name + "__mask" in z.files else None
        ps, po = z[name + "__patch_shape"], z[name + "__patch_overlap"]
        ps = int(ps) if ps.ndim == 0 else tuple(int(v) for v in ps)
        po = int(po) if po.ndim == 0 else tuple(int(v) for v in po)
        common = dict(patch_shape=ps, patch_overlap=po, mask=mask, recombination="weighted")
        for key, call in (("hybrid", lambda: pd.hybrid_pca(data, noise_std=sm32, **common)),
                          ("optimal", lambda: pd.optimal_thresholding(
                              data, loss="fro", noise_std=sm32, **common))):
            out, w, _, _ = call()
            ref, ref_w = z[f"{name}__{key}__out"], z[f"{name}__{key}__w"]
            assert np.array_equal(np.isnan(out), np.isnan(ref))
            e, ew = rel_l2(out, ref), rel_l2(w, ref_w)
            print(f"[gpu-map32] {name} {key}: out {e:.2e} weights {ew:.2e}", flush=True)
            assert e <= 1e-5 and ew <= 1e-6
