"""Seeded synthetic 4-D (space x time) workloads.

Every generator is deterministic (fixed seeds) and documented with the exact
recipe, so a result can be regenerated anywhere.  Shapes are ``(X, Y, Z, T)``
in C order with time contiguous, the layout the denoisers consume.

* :func:`config1`  float32 32x32x16x50, three Gaussian-blob x sinusoid terms
* :func:`config2`  float32 80x80x48x300 ellipsoid "brain", rank-8 AR(1) signal
* :func:`config3`  complex64 64x64x40x200, rank-5 magnitude x smooth phase
* :func:`config4`  eight mixed-shape float32 volumes with per-volume masks
* :func:`config5`  float32 64x64x32x120 rank-6 volume and its clean signal
* :func:`golden_volumes` tiny volumes used for reference fixtures
"""

import numpy as np


def _ellipsoid(shape, semi_axes):
    """Boolean ellipsoid centred at ((S-1)/2) with the given semi-axes."""
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    r2 = np.zeros(shape)
    for g, s, a in zip(grids, shape, semi_axes):
        r2 += ((g - (s - 1) / 2.0) / a) ** 2
    return r2 <= 1.0


def _ar1(rng, n_series, T, rho=0.9):
    """Unit-variance AR(1) time courses, shape (n_series, T)."""
    out = np.empty((n_series, T))
    out[:, 0] = rng.standard_normal(n_series)
    innov = rng.standard_normal((n_series, T)) * np.sqrt(1.0 - rho * rho)
    for t in range(1, T):
        out[:, t] = rho * out[:, t - 1] + innov[:, t]
    return out


def _smooth_maps(rng, n_maps, shape, sigma=3.0):
    """Gaussian-filtered N(0,1) maps, each rescaled to unit standard deviation."""
    from scipy.ndimage import gaussian_filter

    maps = np.empty((n_maps,) + tuple(shape))
    for k in range(n_maps):
        m = gaussian_filter(rng.standard_normal(shape), sigma)
        maps[k] = m / m.std()
    return maps


def config1():
    """float32 32x32x16x50 (seed 0).

    Baseline 500 plus sum_k 100 * G_k(x) * sin(2 pi t / P_k), P = (7, 13, 23),
    G_k an isotropic Gaussian blob (sigma 4 voxels) centred at (8,8,4),
    (16,24,8), (24,12,12); i.i.d. N(0, 20^2) noise.
    """
    rng = np.random.default_rng(0)
    shape, T = (32, 32, 16), 50
    x, y, z = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    t = np.arange(T, dtype=np.float64)
    sig = np.full(shape + (T,), 500.0)
    for (cx, cy, cz), P in zip(((8, 8, 4), (16, 24, 8), (24, 12, 12)), (7, 13, 23)):
        blob = np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * 4.0**2))
        sig += 100.0 * blob[..., None] * np.sin(2 * np.pi * t / P)[None, None, None, :]
    noisy = sig + 20.0 * rng.standard_normal(sig.shape)
    return noisy.astype(np.float32)


def config2(shape=(80, 80, 48), T=300, return_clean=False, seed=2):
    """float32 80x80x48x300 MP-PCA workload with an ellipsoid mask.

    Mask: ellipsoid, semi-axes (36, 36, 20), centred at (S-1)/2.  Inside the
    mask: 800 + sum_{k<8} 40 * S_k(x) a_k(t), S_k = gaussian_filter(N(0,1), 3)
    rescaled to unit std, a_k unit-variance AR(1) with rho = 0.9.  Outside: 0.
    Additive i.i.d. N(0, 30^2) noise everywhere.  Returns ``(data, mask)``
    (and the clean signal when ``return_clean``).
    """
    rng = np.random.default_rng(seed)
    mask = _ellipsoid(shape, (36, 36, 20) if shape == (80, 80, 48) else
                      tuple(0.45 * s for s in shape))
    maps = _smooth_maps(rng, 8, shape).astype(np.float32)
    tcs = _ar1(rng, 8, T).astype(np.float32)
    clean = np.tensordot(maps, 40.0 * tcs, axes=([0], [0])).astype(np.float32)
    clean += np.float32(800.0)
    clean *= mask[..., None]
    noisy = clean + rng.standard_normal(clean.shape, dtype=np.float32) * np.float32(30.0)
    if return_clean:
        return noisy, mask, clean
    return noisy, mask


def config3(shape=(64, 64, 40), T=200, seed=3):
    """complex64 64x64x40x200 NORDIC workload.

    Magnitude: 1000 + sum_{k<5} 60 * S_k(x) a_k(t) (same spatial / AR(1) recipe).
    Phase: phi(x, t) = pi/2 * S_phase(x) + 0.5 * t / (T - 1), S_phase a smooth
    unit-std map.  Complex Gaussian noise, sigma 25 per channel.
    """
    rng = np.random.default_rng(seed)
    maps = _smooth_maps(rng, 6, shape).astype(np.float32)
    tcs = _ar1(rng, 5, T).astype(np.float32)
    mag = np.tensordot(maps[:5], 60.0 * tcs, axes=([0], [0])) + 1000.0
    phi = (np.pi / 2) * maps[5][..., None] + 0.5 * np.arange(T) / (T - 1)
    clean = (mag * np.exp(1j * phi)).astype(np.complex64)
    noise = rng.standard_normal(clean.shape, dtype=np.float32) + 1j * rng.standard_normal(
        clean.shape, dtype=np.float32
    )
    return (clean + np.complex64(25.0) * noise.astype(np.complex64)).astype(np.complex64)


def config5(shape=(64, 64, 32), T=120, seed=5):
    """float32 64x64x32x120: 500 + rank-6 (60 * S_k a_k) signal + N(0, 40^2).

    Returns ``(noisy, clean)``.
    """
    rng = np.random.default_rng(seed)
    maps = _smooth_maps(rng, 6, shape).astype(np.float32)
    tcs = _ar1(rng, 6, T).astype(np.float32)
    clean = (np.tensordot(maps, 60.0 * tcs, axes=([0], [0])) + 500.0).astype(np.float32)
    noisy = clean + rng.standard_normal(clean.shape, dtype=np.float32) * np.float32(40.0)
    return noisy.astype(np.float32), clean


def config4(seed=4, n_volumes=8, scale=1.0):
    """Eight float32 volumes of mixed shape / length / patch size / method.

    Spatial shape: X, Y ~ U{48..96}, Z ~ U{48..64} (times ``scale``),
    T ~ U{80, 120, 200, 300, 400}; per-volume ellipsoid mask (semi-axes 0.45 S);
    patch p ~ U{3, 5, 7}, overlap p - 1.  Methods cycle through mp-pca,
    hybrid-pca (sigma map 20 (1 + 0.5 rho / R)), optimal thresholding (only
    where T <= p^3, otherwise mp-pca) and raw SVT (tau = 20 (sqrt(M)+sqrt(T))).
    Signal: 600 + rank-4 smooth x AR(1), noise N(0, 20^2).
    """
    rng = np.random.default_rng(seed)
    vols = []
    methods = ["mppca", "hybrid", "optimal", "raw"]
    for v in range(n_volumes):
        X = int(rng.integers(48, 97) * scale)
        Y = int(rng.integers(48, 97) * scale)
        Z = int(rng.integers(48, 65) * scale)
        T = int(rng.choice([80, 120, 200, 300, 400]))
        p = int(rng.choice([3, 5, 7]))
        shape = (X, Y, Z)
        mask = _ellipsoid(shape, tuple(0.45 * s for s in shape))
        maps = _smooth_maps(rng, 4, shape).astype(np.float32)
        tcs = _ar1(rng, 4, T).astype(np.float32)
        clean = (np.tensordot(maps, 40.0 * tcs, axes=([0], [0])) + 600.0).astype(np.float32)
        clean *= mask[..., None]
        noisy = clean + 20.0 * rng.standard_normal(clean.shape, dtype=np.float32)
        method = methods[v % 4]
        if method == "optimal" and T > p**3:
            method = "mppca"
        grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        rad = np.sqrt(sum((g - (s - 1) / 2.0) ** 2 for g, s in zip(grids, shape)))
        sigma_map = 20.0 * (1.0 + 0.5 * rad / rad.max())
        vols.append(dict(data=noisy.astype(np.float32), mask=mask, patch=p, method=method,
                         sigma_map=sigma_map,
                         tau=20.0 * (np.sqrt(p**3) + np.sqrt(T))))
    return vols


def config4_calls(vols, max_n=512):
    """Functional-API calls for the batched workload (``config4()`` volumes):
    mp_pca / hybrid_pca (noise_std = the sigma map) / optimal_thresholding /
    raw_svt (threshold = tau), patch p, overlap p - 1, the volume's mask.
    Volumes whose Gram size min(p^3, T) exceeds ``max_n`` are returned in
    ``skipped`` (index, reason) instead."""
    calls, skipped = [], []
    names = {"mppca": "mp_pca", "hybrid": "hybrid_pca", "optimal": "optimal_thresholding",
             "raw": "raw_svt"}
    for i, v in enumerate(vols):
        p, T = v["patch"], v["data"].shape[-1]
        n = min(p**3, T)
        if n > max_n:
            skipped.append((i, f"Gram size min({p}^3, {T}) = {n} > {max_n}"))
            continue
        kw = {"mask": v["mask"]}
        if v["method"] == "hybrid":
            kw["noise_std"] = v["sigma_map"]
        elif v["method"] == "raw":
            kw["threshold"] = float(v["tau"])
        calls.append((names[v["method"]], (v["data"], p, p - 1), kw))
    return calls, skipped


def golden_volumes():
    """Tiny seeded volumes for reference fixtures (each runs in seconds on CPU)."""
    rng = np.random.default_rng(11)

    def lowrank(shape, T, rank, amp, base, sigma, dtype=np.float32, cplx=False):
        maps = rng.standard_normal((rank,) + shape)
        tcs = rng.standard_normal((rank, T))
        clean = base + amp * np.tensordot(maps, tcs, axes=([0], [0]))
        noise = sigma * rng.standard_normal(shape + (T,))
        if cplx:
            phase = np.exp(1j * 0.3 * rng.standard_normal(shape))[..., None]
            noise = noise + 1j * sigma * rng.standard_normal(shape + (T,))
            return (clean * phase + noise).astype(dtype)
        return (clean + noise).astype(dtype)

    vols = {}
    # T-side (M = 27 > T = 24) with a 3-D mask selecting a subset of patches.
    shape = (10, 10, 8)
    mask = _ellipsoid(shape, (4.5, 4.5, 3.5))
    vols["tside_mask"] = dict(
        data=lowrank(shape, 24, 3, 6.0, 100.0, 4.0), mask=mask,
        patch_shape=3, patch_overlap=2, sigma=4.0,
        sigma_map=4.0 * (1 + 0.1 * rng.random(shape)), raw_tau=4.0 * (np.sqrt(27) + np.sqrt(24)),
    )
    # M-side (M = 27 < T = 40), anisotropic patch with uncovered tails (NaN).
    shape = (12, 10, 6)
    vols["mside_aniso"] = dict(
        data=lowrank(shape, 40, 2, 5.0, 50.0, 3.0), mask=None,
        patch_shape=(3, 3, 3), patch_overlap=(1, 2, 1), sigma=3.0,
        sigma_map=3.0 * (1 + 0.1 * rng.random(shape)), raw_tau=3.0 * (np.sqrt(27) + np.sqrt(40)),
        optimal_ok=False,
    )
    # 2-D + time, like the reference's own tests (patch 6, overlap 5 -> 36 x 20).
    shape = (16, 14)
    vols["2d_p6"] = dict(
        data=lowrank(shape, 20, 2, 4.0, 30.0, 1.0), mask=None,
        patch_shape=6, patch_overlap=5, sigma=1.0,
        sigma_map=1.0 * (1 + 0.2 * rng.random(shape)), raw_tau=1.0 * (6.0 + np.sqrt(20)),
    )
    # complex64, T-side.
    shape = (8, 8, 6)
    vols["complex_tside"] = dict(
        data=lowrank(shape, 20, 2, 5.0, 40.0, 2.0, dtype=np.complex64, cplx=True), mask=None,
        patch_shape=3, patch_overlap=2, sigma=2.0,
        sigma_map=2.0 * (1 + 0.1 * rng.random(shape)),
        raw_tau=2.0 * np.sqrt(2) * (np.sqrt(27) + np.sqrt(20)),
    )
    # float64 input, M-side for SVD methods excluded (T > M raises for optimal).
    shape = (8, 8, 6)
    vols["f64_mside"] = dict(
        data=lowrank(shape, 32, 2, 5.0, 20.0, 2.0, dtype=np.float64), mask=None,
        patch_shape=3, patch_overlap=2, sigma=2.0,
        sigma_map=2.0 * (1 + 0.1 * rng.random(shape)), raw_tau=2.0 * (np.sqrt(27) + np.sqrt(32)),
        optimal_ok=False,
    )
    # 4-D mask and mask_threshold 70.
    shape = (9, 9, 7)
    m4 = np.broadcast_to(_ellipsoid(shape, (4.0, 4.0, 3.0))[..., None], shape + (22,)).copy()
    m4[..., :3] = False
    vols["mask4d_thr70"] = dict(
        data=lowrank(shape, 22, 2, 5.0, 60.0, 3.0), mask=m4, mask_threshold=70,
        patch_shape=3, patch_overlap=2, sigma=3.0,
        sigma_map=3.0 * (1 + 0.1 * rng.random(shape)), raw_tau=3.0 * (np.sqrt(27) + np.sqrt(22)),
    )
    return vols


def golden_volumes_small_m():
    """Tiny seeded volumes whose patches hold fewer than 10 voxels (M < 10):
    the eigenvalue rules then see fewer than 10 nonzero Gram eigenvalues.

    * ``2d_p3``: 2-D 14x12 x 30 frames, patch 3, overlap 2 (M = 9);
    * ``3d_p2``: 3-D 8x8x6 x 24 frames, patch 2, overlap 1 (M = 8).
    Rank-2 signal plus Gaussian noise; optimal shrinkage is excluded (the
    reference raises when the patch has fewer voxels than frames)."""
    rng = np.random.default_rng(12)

    def lowrank(shape, T, rank, amp, base, sigma):
        maps = rng.standard_normal((rank,) + shape)
        tcs = rng.standard_normal((rank, T))
        clean = base + amp * np.tensordot(maps, tcs, axes=([0], [0]))
        return (clean + sigma * rng.standard_normal(shape + (T,))).astype(np.float32)

    vols = {}
    shape = (14, 12)
    vols["2d_p3"] = dict(
        data=lowrank(shape, 30, 2, 5.0, 40.0, 2.0), mask=None,
        patch_shape=3, patch_overlap=2, sigma=2.0,
        sigma_map=2.0 * (1 + 0.1 * rng.random(shape)), raw_tau=2.0 * (3.0 + np.sqrt(30)),
        optimal_ok=False,
    )
    shape = (8, 8, 6)
    vols["3d_p2"] = dict(
        data=lowrank(shape, 24, 2, 5.0, 40.0, 2.0), mask=None,
        patch_shape=2, patch_overlap=1, sigma=2.0,
        sigma_map=2.0 * (1 + 0.1 * rng.random(shape)), raw_tau=2.0 * (np.sqrt(8) + np.sqrt(24)),
        optimal_ok=False,
    )
    return vols
