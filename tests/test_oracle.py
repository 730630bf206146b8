"""Pin the CPU oracle: golden fixtures from the reference + closed forms.

These run on CPU only (no GPU marker).
"""

import glob
import os

import numpy as np
import pytest

from oracle import cpu_ref as O

from .golden.cases import run_table

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E2E = sorted(glob.glob(os.path.join(GOLDEN, "e2e_*.npz")))


def _spec(z):
    return {k: z[k] for k in ("sigma", "sigma_map", "raw_tau", "optimal_ok")}


def _assert_close(got, want, dtype, what):
    assert got.dtype == want.dtype, what
    assert got.shape == want.shape, what
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    assert np.array_equal(nan_g, nan_w), f"{what}: NaN pattern differs"
    tol = 2e-5 if np.dtype(dtype).itemsize <= 8 and dtype != np.float64 else 1e-12
    g, w = got[~nan_g], want[~nan_w]
    scale = max(np.abs(w).max(), 1e-30) if w.size else 1.0
    assert np.abs(g - w).max() <= tol * scale, f"{what}: max abs diff {np.abs(g - w).max()}"


@pytest.mark.parametrize("path", E2E, ids=[os.path.basename(p)[4:-4] for p in E2E])
def test_oracle_e2e_matches_reference_fixtures(path):
    z = np.load(path)
    data = z["input"]
    mask = z["mask"] if "mask" in z.files else None
    ps = z["patch_shape"]
    po = z["patch_overlap"]
    ps = int(ps) if ps.ndim == 0 else tuple(int(v) for v in ps)
    po = int(po) if po.ndim == 0 else tuple(int(v) for v in po)
    common = dict(patch_shape=ps, patch_overlap=po, mask=mask,
                  mask_threshold=int(z["mask_threshold"]))
    for key, fn_name, kw in run_table(_spec(z)):
        if fn_name == "nordic":
            np.random.seed(0)
        out, w, noise, rank = getattr(O, fn_name)(data, **common, **kw)
        _assert_close(out, z[key + "__out"], data.dtype, key + " out")
        _assert_close(w, z[key + "__w"], np.float32, key + " weights")
        assert not noise.any() and not rank.any()


def test_oracle_per_patch_fixtures():
    z = np.load(os.path.join(GOLDEN, "patches_cfg1.npz"))
    for j, X in enumerate(z["casorati"]):
        _, a, _, _ = O.eig_analysis(X.copy())
        np.testing.assert_allclose(a, z["eig_top10"][j], rtol=1e-5)
        P, r, v = O.mppca_patch(X.copy())
        assert r == z["mppca_rank"][j]
        np.testing.assert_allclose(v, z["mppca_noisevar"][j], rtol=1e-5)
        np.testing.assert_allclose(P, z["mppca_patch"][j], rtol=1e-5, atol=1e-3)
        _, s, _, _ = O.svd_analysis(X.copy())
        np.testing.assert_allclose(s, z["svd_s"][j], rtol=1e-5, atol=1e-4)


def test_oracle_cfg1_full_run():
    from patch_denoise import workloads

    z = np.load(os.path.join(GOLDEN, "cfg1_mppca.npz"))
    data = workloads.config1()
    out, w, _, _ = O.mp_pca(data, patch_shape=5, patch_overlap=2)
    assert int(np.isnan(out).sum()) == int(z["out_nan_count"])
    flat = z["sample_index"]
    np.testing.assert_allclose(out.reshape(-1)[flat], z["sample_out"], rtol=1e-5, equal_nan=True)
    np.testing.assert_allclose(w.reshape(-1)[flat], z["sample_w"], rtol=1e-6)
    ps, po = O.patch_params(5, 2, data.shape)
    src = O.StridedPatches(data, ps, po)
    for i in range(0, src.n_patches, 7):
        _, r, v = O.mppca_patch(src.get(i).reshape(125, 50))
        assert r == z["ranks"][i]


def test_mp_median_fixture_and_closed_form():
    z = np.load(os.path.join(GOLDEN, "mp_median.npz"))
    for i, b in enumerate(z["beta"]):
        for j, e in enumerate(z["eps"]):
            assert O.mp_median(b, eps=e) == pytest.approx(z["median"][i, j], rel=1e-12, abs=1e-14)
    # Independent pin: root of the MP CDF (closed-form pdf, adaptive quadrature).
    from scipy.integrate import quad
    from scipy.optimize import brentq

    for b in (0.1, 0.4, 0.9):
        lo, hi = (1 - np.sqrt(b)) ** 2, (1 + np.sqrt(b)) ** 2

        def cdf(x):
            f = lambda t: np.sqrt(max((hi - t) * (t - lo), 0.0)) / (2 * np.pi * t * b)
            return quad(f, lo, x, epsabs=1e-13, epsrel=1e-13, limit=200)[0]

        root = brentq(lambda x: cdf(x) - 0.5, lo, hi, xtol=1e-14)
        assert abs(O.mp_median(b, eps=1e-7) - root) < 2e-7
    with pytest.raises(ValueError):
        O.mp_median(1.5)


def test_nordic_tau_fixture():
    z = np.load(os.path.join(GOLDEN, "nordic_tau.npz"))
    for n, T, it, seed, tau in zip(z["patch_shape"], z["T"], z["n_iter"], z["seed"], z["tau"]):
        np.random.seed(int(seed))
        got = O.nordic_threshold((int(n),), int(T), 1.0, int(it))
        assert got == pytest.approx(float(tau), rel=1e-12)
        # Bai-Yin: the top singular value of an N x T Gaussian is ~ sqrt(N) + sqrt(T).
        assert abs(got - (np.sqrt(n) + np.sqrt(T))) < 0.15 * (np.sqrt(n) + np.sqrt(T))
    with pytest.raises(ValueError):
        O.nordic_threshold((27,), 20, "nope", 2)


def test_grid_and_selection_brute_force():
    rng = np.random.default_rng(0)
    shape = (7, 6, 5, 4)
    mask = rng.random(shape[:-1]) > 0.4
    for ps, po in (((3, 3, 3, 4), (2, 2, 2, 4)), ((3, 2, 4, 4), (1, 0, 2, 4)), ((2, 2, 2, 4), (0, 0, 0, 4))):
        sel, _ = O.select_patches(mask, shape, ps, po, 50)
        step = [p - o for p, o in zip(ps, po)]
        grid = [((s - p) // st + 1) if p < s else 1 for s, p, st in zip(shape, ps, step)]
        want = []
        lin = 0
        for gx in range(grid[0]):
            for gy in range(grid[1]):
                for gz in range(grid[2]):
                    for gt in range(grid[3]):
                        x0, y0, z0 = gx * step[0], gy * step[1], gz * step[2]
                        cnt = mask[x0:x0 + ps[0], y0:y0 + ps[1], z0:z0 + ps[2]].sum() * ps[3]
                        if 100 * cnt / np.prod(ps) > 50:
                            want.append(lin)
                        lin += 1
        assert list(sel) == want
    with pytest.raises(ValueError):
        O.select_patches(np.ones((3, 3), bool), shape, (3, 3, 3, 4), (2, 2, 2, 4), 50)
    with pytest.raises(ValueError):
        O.StridedPatches(np.zeros(shape), (3, 3, 3, 4), (4, 2, 2, 4))


def test_projection_identities():
    """eig/svd analysis+synthesis: full rank reproduces X; rank-r is a projection."""
    rng = np.random.default_rng(1)
    X = rng.standard_normal((40, 12)) * 3 + 5
    U, s, Vt, m = O.svd_analysis(X.copy())
    np.testing.assert_allclose(O.svd_synthesis(U, s, Vt, m, len(s)), X, atol=1e-10)
    Xc, a, W, m2 = O.eig_analysis(X.copy())
    np.testing.assert_allclose(np.sort(s**2 / 39)[-10:], a, rtol=1e-9)
    P = O.eig_synthesis(Xc, W.copy(), m2, 4) - m2
    Q = U[:, :4] @ U[:, :4].T @ (X - m)
    np.testing.assert_allclose(P, Q, atol=1e-9)


def test_shrinkers_closed_forms():
    b = 0.25
    # ope: x(y) inverts the spiked-model map y(x) = sqrt((x + 1/x)(x + b/x))
    # above the bulk edge 1 + sqrt(b); inside the bulk the reference yields NaN
    # (0 * NaN), which the caller later zeroes.
    y = np.array([0.5, 1.49, 1.5, 2.0, 4.0])
    with np.errstate(invalid="ignore"):
        x = O.shrink_ope(y, b)
    assert np.isnan(x[0]) and np.isnan(x[1])
    xs = x[2:]
    np.testing.assert_allclose(np.sqrt((xs + 1 / xs) * (xs + b / xs)), y[2:], rtol=1e-12)
    # fro leaks below the lower edge 1 - sqrt(b) = 0.5 (no gate) and is 0 inside the bulk.
    with np.errstate(divide="ignore"):
        f = O.shrink_fro(np.array([0.3, 0.5, 1.0, 1.5, 3.0]), b)
    assert f[0] > 0 and f[2] == 0 and f[1] == 0 and f[3] == 0 and f[4] > 0
    np.testing.assert_allclose(f[4], np.sqrt(((9 - b - 1) ** 2 - 4 * b) / 3.0))
    # nuc: zero until x^4 - sqrt(b) x y - b > 0; NaN inside the bulk (0/0).
    with np.errstate(invalid="ignore", divide="ignore"):
        n = O.shrink_nuc(np.array([1.0, 3.0]), b)
    assert np.isnan(n[0]) and n[1] > 0


def test_adaptive_parameters_match_reference_fixtures():
    """(gamma, tau) of the oracle's SURE / GSURE / QUT search against the
    reference's own on four Casorati matrices (tests/golden/adaptive.npz)."""
    import warnings

    z = np.load(os.path.join(GOLDEN, "adaptive.npz"))
    gam = z["gamma_grid"]
    sd = float(z["stdest"])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for k in range(4):
            X, s = z[f"patch{k}_X"], z[f"patch{k}_s"]
            for m in ("sure", "gsure"):
                g, t = O.adaptive_params(X.shape, s, sd, m, gam, None, 0)
                assert g == z[f"patch{k}_{m}_gamma"]
                np.testing.assert_allclose(t, z[f"patch{k}_{m}_tau"], rtol=1e-7)
            for label, g0 in (("qutg2", 2.0), ("qut", None)):
                np.random.seed(100 + k)
                g, t = O.adaptive_params(X.shape, s, sd, "qut", g0, None, 60)
                np.testing.assert_allclose(np.ravel(g), z[f"patch{k}_{label}_gamma"], rtol=1e-7)
                np.testing.assert_allclose(t, z[f"patch{k}_{label}_tau"], rtol=1e-12)


def test_adaptive_end_to_end_matches_reference_fixtures():
    import warnings

    z = np.load(os.path.join(GOLDEN, "adaptive.npz"))
    gam = z["gamma_grid"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for v in ("v3d", "v2d"):
            data, ps, po = z[f"{v}_in"], int(z[f"{v}_ps"]), int(z[f"{v}_po"])
            for method, g0, recomb in (("SURE", gam, "weighted"), ("GSURE", gam, "average"),
                                       ("qut", 2.0, "weighted")):
                np.random.seed(7)
                out, w, _, _ = O.adaptive_thresholding(data, ps, po, recombination=recomb,
                                                       method=method, nbsim=40, gamma0=g0,
                                                       noise_std=1.5)
                want = z[f"{v}_{method.lower()}_out"]
                assert np.array_equal(np.isnan(out), np.isnan(want))
                fin = np.isfinite(want)
                np.testing.assert_allclose(out[fin], want[fin], rtol=0, atol=1e-5 * np.abs(want[fin]).max())
                np.testing.assert_array_equal(w, z[f"{v}_{method.lower()}_w"])


def test_adaptive_cost_limits_closed_form():
    """SURE at tau -> 0 is the unbiased risk of X itself, sigma^2 n p (the
    divergence of the identity is n p: k^2 + k |n - p| = k max(n, p) with
    k = min(n, p)); at tau -> infinity it is the risk of 0, |X|^2 - sigma^2 n p.
    A dropped pair term, a wrong |n - p| factor or sign breaks one of them."""
    rng = np.random.default_rng(8)
    for nr, nc in ((12, 7), (7, 12), (9, 9)):
        s = np.sort(rng.uniform(1.0, 10.0, min(nr, nc)))[::-1].astype(np.float64)
        sig = 0.7
        lo = O.atn_cost((nr, nc), "sure", s, 2.0, sig, np.log(1e-9))
        hi = O.atn_cost((nr, nc), "sure", s, 2.0, sig, np.log(1e9))
        np.testing.assert_allclose(lo, sig**2 * nr * nc, rtol=1e-5)
        np.testing.assert_allclose(hi, np.sum(s**2) - sig**2 * nr * nc, rtol=1e-9)
        # GSURE at tau -> infinity: rss / 1
        np.testing.assert_allclose(O.atn_cost((nr, nc), "gsure", s, 2.0, sig, np.log(1e9)),
                                   np.sum(s**2), rtol=1e-9)
        # shrinker: hard zero below tau, s - tau at gamma = 1 above
        sh = O.atn_shrink(s, 1.0, 4.0)
        np.testing.assert_allclose(sh, np.maximum(s - 4.0, 0), rtol=1e-12)
