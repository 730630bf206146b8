"""GPU parity for Gram sizes n = min(patch voxels, frames) above 128 (and a
complex64 Gram of 100, on the same register-tile Householder kernel).

Patch shapes of the full-size workloads whose Gram exceeds 128 -- 11^3 x 200
complex NORDIC patches, 7^3-voxel patches of 200 to 400 frames, 6^3 x 300,
15^2 x 300 in 2-D, and Grams above 512 up to 11^3 = 1331 voxels over 1400
frames -- on volumes just a few patches wide, so that the oracle
(a NumPy restatement of the reference, SVD / eigh per patch) finishes in
seconds.  Both Gram sides are covered: T-side (frames < voxels, G = Xc^H Xc)
and M-side (voxels <= frames, G = Xc Xc^H).

Tolerances as for the small Grams: relative L2 over finite entries with the
NaN pattern required to match; 1e-5 for float32 / complex64 data, 5e-7 for
float64 (the float32 weight divisor differs by float32 ulps with the
summation order); weights 1e-6.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(got, want):
    fin = np.isfinite(want)
    d = np.asarray(got, np.complex128)[fin] - np.asarray(want, np.complex128)[fin]
    den = np.linalg.norm(np.asarray(want, np.complex128)[fin])
    return float(np.linalg.norm(d) / max(den, 1e-300))


def _lowrank(shape, T, rank, seed, dtype=np.float32, sigma=30.0, base=800.0, amp=60.0):
    rng = np.random.default_rng(seed)
    maps = rng.standard_normal((rank,) + shape)
    tcs = rng.standard_normal((rank, T))
    clean = base + amp * np.tensordot(maps, tcs, axes=([0], [0]))
    return (clean + sigma * rng.standard_normal(shape + (T,))).astype(dtype)


def _lowrank_c(shape, T, rank, seed, sigma=30.0, base=800.0, amp=60.0):
    """complex64 low-rank volume of the same dynamic range as ``_lowrank``
    (complex maps and time courses, circular complex noise)."""
    rng = np.random.default_rng(seed)
    cn = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    clean = base + amp * np.tensordot(cn(rank, *shape), cn(rank, T), axes=([0], [0])) / 2.0
    return (clean + sigma / np.sqrt(2.0) * cn(*shape, T)).astype(np.complex64)


def _sigma_map(shape, s0=30.0):
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    rad = np.sqrt(sum((g - (s - 1) / 2.0) ** 2 for g, s in zip(grids, shape)))
    return s0 * (1.0 + 0.5 * rad / max(rad.max(), 1e-9))


def _cases():
    from patch_denoise import workloads

    c = []
    # T-side, n = 200 (7^3 = 343 voxels > 200 frames), MP-PCA float32
    c.append(("mppca_f32_p7_T200", "mp_pca", _lowrank((9, 9, 8), 200, 6, 1), (7, 6), {}))
    # real float32 Grams of 129 .. 208 (four-wavefront register Householder):
    # the smallest size, sizes off the 16- and 8-row tile grid, the largest,
    # and an M-side Hybrid-PCA Gram
    c.append(("mppca_f32_p7_T131", "mp_pca", _lowrank((8, 8, 8), 131, 4, 12), (7, 6), {}))
    c.append(("mppca_f32_p7_T150", "mp_pca", _lowrank((9, 9, 8), 150, 5, 10), (7, 6), {}))
    c.append(("mppca_f32_p7_T180", "mp_pca", _lowrank((8, 9, 8), 180, 5, 14), (7, 6), {}))
    c.append(("mppca_f32_p7_T208", "mp_pca", _lowrank((8, 8, 9), 208, 5, 13), (7, 6), {}))
    d = _lowrank((8, 8, 6), 300, 4, 11)
    c.append(("hybrid_f32_p6x6x4_T300", "hybrid_pca", d, ((6, 6, 4), (5, 5, 3)),
              {"noise_std": _sigma_map(d.shape[:-1])}))
    # the same reduction in FP64 (Grams of 65 .. 208): raw SVT and optimal
    # shrinkage on float32 data (FP64 eigen stage by default), float64 data
    c.append(("raw_f32_p7_T200", "raw_svt", _lowrank((9, 9, 8), 200, 5, 15), (7, 6),
              {"threshold": 30.0 * (np.sqrt(343) + np.sqrt(200)) * 1.02}))
    c.append(("optimal_fro_f32_p7_T150", "optimal_thresholding", _lowrank((8, 8, 9), 150, 5, 16),
              (7, 6), {"loss": "fro"}))
    c.append(("mppca_f64_p7_T170", "mp_pca", _lowrank((8, 8, 8), 170, 4, 17, dtype=np.float64),
              (7, 6), {}))
    # complex64 MP-PCA (complex FP32 eigen stage, Hermitian reduction): n = 100
    # (packed reflector rows) and n = 200
    c.append(("mppca_c64_p5_T100", "mp_pca", _lowrank_c((8, 8, 7), 100, 4, 18), (5, 4), {}))
    c.append(("mppca_c64_p7_T200", "mp_pca", _lowrank_c((9, 8, 8), 200, 5, 19), (7, 6), {}))
    c.append(("hybrid_c64_p7_T160", "hybrid_pca", _lowrank_c((8, 8, 8), 160, 4, 20), (7, 6),
              {"noise_std": 30.0}))
    # M-side, n = 216 (6^3 voxels < 260 frames), MP-PCA float64
    c.append(("mppca_f64_p6_T260", "mp_pca",
              _lowrank((8, 8, 7), 260, 5, 2, dtype=np.float64), (6, 5), {}))
    # M-side, n = 343, Hybrid-PCA with a per-voxel sigma map
    d = _lowrank((8, 8, 8), 400, 4, 3)
    c.append(("hybrid_f32_p7_T400", "hybrid_pca", d, (7, 6),
              {"noise_std": _sigma_map(d.shape[:-1])}))
    # M-side, n = 216, raw SVT, spatial mask, average recombination
    d = _lowrank((8, 8, 7), 300, 5, 4)
    mask = np.ones(d.shape[:-1], bool)
    mask[0, :, :] = False
    c.append(("raw_f32_p6_T300_mask_avg", "raw_svt", d, (6, 5),
              {"threshold": 30.0 * (np.sqrt(216) + np.sqrt(300)) * 1.05, "mask": mask,
               "recombination": "average"}))
    # T-side, n = 300 (beta = 300 / 343), optimal shrinkage, Frobenius and nuclear
    d = _lowrank((8, 8, 8), 300, 6, 5)
    c.append(("optimal_fro_f32_p7_T300", "optimal_thresholding", d, (7, 6), {"loss": "fro"}))
    c.append(("optimal_nuc_f32_p7_T300_center", "optimal_thresholding", d, (7, 6),
              {"loss": "nuc", "recombination": "center"}))
    # 2-D: 15^2 = 225 voxels < 300 frames -> M-side n = 225, MP-PCA
    c.append(("mppca_2d_f32_p15_T300", "mp_pca", _lowrank((17, 17), 300, 4, 6), (15, 14), {}))
    # complex64 NORDIC on 11^3 x 200 patches (T-side, n = 200)
    d = workloads.config3(shape=(12, 12, 11), T=200)
    c.append(("nordic_c64_p11_T200", "nordic", d, (11, 10),
              {"noise_std": 25.0, "n_iter_threshold": 1}))
    # Grams above 512 (one eigenvector component per batch, the Householder
    # reduction mostly on the matrix in global memory):
    # M-side n = 648 (9x9x8 voxels < 700 frames), MP-PCA float32
    c.append(("mppca_f32_p9x9x8_T700", "mp_pca", _lowrank((10, 10, 9), 700, 5, 7),
              ((9, 9, 8), (8, 8, 7)), {}))
    # T-side n = 520 (11^3 voxels > 520 frames), complex64 raw SVT, FP64 eigen
    # stage (NORDIC's threshold draw alone, a 692,120 x 520 SVD, would take
    # the oracle a minute here)
    d = workloads.config3(shape=(12, 11, 11), T=520)
    c.append(("raw_c64_p11_T520", "raw_svt", d, (11, 10),
              {"threshold": 25.0 * np.sqrt(2) * (np.sqrt(1331) + np.sqrt(520))}))
    # M-side n = 1331 (11^3 voxels < 1400 frames), MP-PCA float32: the
    # reference command line's default patch with a long series
    c.append(("mppca_f32_p11_T1400", "mp_pca", _lowrank((12, 11, 11), 1400, 5, 8), (11, 10), {}))
    # T-side n = 700, optimal shrinkage (Frobenius), float32
    c.append(("optimal_fro_f32_p11_T700", "optimal_thresholding",
              _lowrank((12, 11, 11), 700, 6, 9), (11, 10), {"loss": "fro"}))
    return c


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_large_gram_matches_oracle(case):
    import patch_denoise as pd
    from oracle import cpu_ref as O

    name, fn, data, (ps, po), kw = case
    vox = int(np.prod(ps)) if isinstance(ps, tuple) else ps ** (data.ndim - 1)
    n = min(vox, data.shape[-1])
    assert n > 64
    print(f"[large-gram] {name}: shape {data.shape} {data.dtype} patch {ps} overlap {po} n {n}",
          flush=True)
    np.random.seed(0)
    out, w, _, _ = getattr(pd, fn)(data, ps, po, **kw)
    np.random.seed(0)
    want, want_w, _, _ = getattr(O, fn)(data, ps, po, **kw)
    assert out.dtype == want.dtype and out.shape == want.shape
    assert np.array_equal(np.isnan(out), np.isnan(want)), "NaN pattern"
    e = rel_l2(out, want)
    ew = rel_l2(w, want_w)
    tol = 5e-7 if data.dtype in (np.float64, np.complex128) else 1e-5
    print(f"[large-gram]   rel-L2 out {e:.3g} weights {ew:.3g} (tolerance {tol:g})", flush=True)
    assert e <= tol and ew <= 1e-6


@pytest.mark.parametrize("which", ["nordic_c64_p11_T200", "mppca_f32_p7_T200", "mppca_2d_f32_p15_T300"])
def test_large_gram_split_bf16(which):
    """precision="bf16x3": the large Gram from split-BF16 three-product MFMA
    (complex: four real products) with FP32 accumulation; the eigen stage in
    FP32.  Bound: rel-L2 <= 1e-3 against the oracle (the split drops the
    lo x lo term, ~2^-16 relative)."""
    import patch_denoise as pd
    from oracle import cpu_ref as O

    name, fn, data, (ps, po), kw = next(c for c in CASES if c[0] == which)
    np.random.seed(0)
    out, w, _, _ = getattr(pd, fn)(data, ps, po, precision="bf16x3", **kw)
    np.random.seed(0)
    ref32, _, _, _ = getattr(pd, fn)(data, ps, po, precision="fp32", **kw)
    np.random.seed(0)
    want, want_w, _, _ = getattr(O, fn)(data, ps, po, **kw)
    e, e32 = rel_l2(out, want), rel_l2(out, ref32)
    print(f"[large-gram] {name} bf16x3: rel-L2 vs oracle {e:.3g}, vs fp32 path {e32:.3g}",
          flush=True)
    assert np.array_equal(np.isnan(out), np.isnan(want))
    assert e <= 1e-3 and rel_l2(w, want_w) <= 1e-6
