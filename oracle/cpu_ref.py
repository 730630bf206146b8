"""CPU oracle for the space-time low-rank denoising path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``patch_denoise``) never imports it and has no
CPU fallback.

It restates, in NumPy/SciPy, the reference's per-patch algorithm and patch
driver with the same LAPACK calls (``eigh(driver="evx")`` top-10 and
``svd`` / gesdd) in the same C-order patch loop, so that on the same inputs it
reproduces the reference's outputs to the rounding of those libraries.  Each
function cites the reference lines it follows (paths relative to the
reference's ``src/patch_denoise``).

Pinning (checked in ``tests/test_oracle.py``):

* every function against the golden fixtures written from the reference
  (``tests/golden/*.npz``: end-to-end outputs for every method x
  recombination, per-patch spectra/ranks, MP median grid, NORDIC tau);
* grid / selection / recombination by brute force on tiny volumes;
* eigen / SVD analysis by the projection and reconstruction identities;
* the MP median against the Marchenko-Pastur CDF root (closed-form pdf);
* the optimal shrinkers against their closed forms at hand-picked points.
"""

import logging

import itertools

import numpy as np
from scipy.integrate import quad
from scipy.linalg import eigh, svd

# --------------------------------------------------------------------------
# Patch geometry (space_time/base.py)
# --------------------------------------------------------------------------


def patch_params(p_shape, p_ovl, data_shape):
    """Normalise patch shape / overlap to per-dimension tuples.

    Follows ``BaseSpaceTimeDenoiser._get_patch_param`` (base.py:359-384): an
    int becomes one value per spatial dim; a spatial-only tuple gets the full
    time length appended to BOTH shape and overlap (one temporal patch); a
    warning is logged when the spatial patch holds fewer voxels than frames.
    """
    out = []
    for p in (p_shape, p_ovl):
        if isinstance(p, list):
            p = tuple(p)
        elif isinstance(p, (int, np.integer)):
            p = (p,) * (len(data_shape) - 1)
        if len(p) == len(data_shape) - 1:
            p = (*p, data_shape[-1])
        out.append(p)
    if np.prod(out[0][:-1]) < data_shape[-1]:
        logging.warning(
            f"the number of voxel in patch ({np.prod(out[0])}) is smaller than the"
            f" last dimension ({data_shape[-1]}), this makes an ill-conditioned"
            "matrix for SVD.",
            stacklevel=2,
        )
    return tuple(out[0]), tuple(out[1])


class StridedPatches:
    """Sliding patch view of an N-D array (base.py:13-150, ``PatchedArray``).

    Grid per dim: ``(S - p) // step + 1`` when ``p < S`` else 1 (base.py:56-63),
    patch clipped to the array (base.py:53), no padding; ``ValueError`` when an
    overlap exceeds the patch (base.py:43-44) or the rank mismatches
    (base.py:46-50).
    """

    def __init__(self, arr, ps, po):
        self.arr = arr
        ps = np.asarray(ps)
        step = ps - np.asarray(po)
        if np.any(step < 0):
            raise ValueError("overlap should be smaller than patch on every dimension.")
        if ps.size != arr.ndim or step.size != arr.ndim:
            raise ValueError("patch shape and step must match the array's dimensions.")
        ps = np.minimum(ps, arr.shape)
        with np.errstate(divide="ignore"):
            grid = tuple(
                int((arr.shape[i] - ps[i]) // step[i] + 1) if ps[i] < arr.shape[i] else 1
                for i in range(arr.ndim)
            )
        strides = tuple(
            arr.strides[i] * step[i] if ps[i] < arr.shape[i] else 0 for i in range(arr.ndim)
        ) + arr.strides
        self.view = np.lib.stride_tricks.as_strided(arr, shape=grid + tuple(ps), strides=strides)
        self.grid = grid
        self.ps = tuple(int(p) for p in ps)

    @property
    def n_patches(self):
        return int(np.prod(self.grid))

    def get(self, idx):
        return self.view[np.unravel_index(idx, self.grid)]


def select_patches(mask, data_shape, ps, po, mask_threshold):
    """Linear indices of processed patches, increasing (base.py:201-227).

    A patch is kept iff ``100 * sum(mask_patch) / mask_patch.size >
    mask_threshold`` (strict).  ``mask=None`` keeps everything; a spatial mask
    is broadcast over time; any other shape raises ``ValueError``.
    Returns ``(selected_indices, full_4d_mask)``.
    """
    if mask is None:
        full = np.full(data_shape, True)
    elif mask.shape == data_shape:
        full = mask
    elif mask.shape == data_shape[:-1]:
        full = np.broadcast_to(mask[..., None], data_shape)
    else:
        raise ValueError(f"Mask shape {mask.shape} is incompatible with input {data_shape}.")
    pm = StridedPatches(full, ps, po)
    keep = np.zeros(pm.n_patches, dtype=bool)
    for i in range(pm.n_patches):
        m = pm.get(i)
        keep[i] = 100 * np.sum(m) / m.size > mask_threshold
    return np.nonzero(keep)[0], full


# --------------------------------------------------------------------------
# Per-patch linear algebra (space_time/utils.py)
# --------------------------------------------------------------------------


def svd_analysis(X):
    """Centred thin SVD (utils.py:58-78): per-frame spatial mean, gesdd."""
    m = np.mean(X, axis=0)
    Xc = X - m
    U, s, Vt = svd(Xc, full_matrices=False)
    return U, s, Vt, m


def svd_synthesis(U, s, Vt, m, r):
    """Rank-r reconstruction ``U_r (s_r Vt_r) + m`` (utils.py:81-99)."""
    return (U[:, :r] @ (s[:r, None] * Vt[:r, :])) + m


def eig_analysis(X, n_top=10):
    """Top-``n_top`` eigenpairs of the centred T x T Gram / (M-1) (utils.py:102-137)."""
    m = np.mean(X, axis=0)
    Xc = X - m
    M = Xc.shape[0]
    w, W = eigh(Xc.conj().T @ Xc / (M - 1), driver="evx",
                subset_by_index=(len(m) - n_top, len(m) - 1))
    return Xc, w, W, m


def eig_synthesis(Xc, W, m, r):
    """Projection on the top-r eigenvectors (utils.py:140-146; W is modified)."""
    W[:, :-r] = 0
    return ((Xc @ W) @ W.conj().T) + m


def mp_median(beta, eps=1e-7):
    """Median of the Marchenko-Pastur law by 5-point bracketing (utils.py:149-204).

    Each round integrates the MP pdf from the lower edge to 5 equispaced
    points with ``scipy.integrate.quad`` and shrinks the bracket to the points
    around 1/2; stops on width <= eps, no change, or 20 rounds; returns the
    bracket midpoint.
    """
    if not (0 <= beta <= 1):
        raise ValueError("Aspect Ratio should be between 0 and 1")
    hi_edge = (1 + np.sqrt(beta)) ** 2
    lo_edge = (1 - np.sqrt(beta)) ** 2

    def pdf(x):
        if hi_edge >= x >= lo_edge:
            return np.sqrt((hi_edge - x) * (x - lo_edge)) / (2 * np.pi * x * beta)
        return 0

    lo, hi = lo_edge, hi_edge
    moved = True
    it = 0
    while moved and (hi - lo) > eps and it < 20:
        moved = False
        pts = np.linspace(lo, hi, 5)
        cdf = np.array([quad(pdf, lo_edge, p)[0] for p in pts])
        if np.any(cdf < 0.5):
            lo = np.max(pts[cdf < 0.5])
            moved = True
        if np.any(cdf > 0.5):
            hi = np.min(pts[cdf > 0.5])
            moved = True
        it += 1
    return (lo + hi) / 2


# --------------------------------------------------------------------------
# Threshold rules (space_time/lowrank.py)
# --------------------------------------------------------------------------


def mppca_patch(X, threshold_scale=1.0):
    """MP-PCA rule on the top-10 eigenvalues (lowrank.py:42-63).

    Note the reference's two different lengths: the first factor uses the
    number of computed eigenvalues (10), the loop uses the eigenvector row
    count (T).  Returns ``(patch, rank, noise_variance)``.
    """
    Xc, a, W, m = eig_analysis(X)
    k = 0
    var = np.mean(a)
    var *= 4 * np.sqrt((len(a) - k + 1) / len(X))
    while k < len(a) and var < a[~k] - a[0]:
        k += 1
        var = np.mean(a[:-k])
        var *= 4 * np.sqrt((len(W) - k + 1) / len(X))
    noise = np.mean(a[: len(a) - k])
    r = np.sum(a > (noise * threshold_scale**2))
    if r == 0:
        out = np.zeros_like(X) + m
    else:
        out = eig_synthesis(Xc, W, m, r)
    return out, r, noise


def hybrid_patch(X, varest):
    """Hybrid-PCA rule (lowrank.py:104-119): drop top eigenvalues while the mean
    of the rest exceeds the prior variance, at most ``len - 2`` of them."""
    Xc, a, W, m = eig_analysis(X)
    k = 0
    var = np.mean(a)
    while var > varest and k < len(a) - 2:
        k += 1
        var = np.mean(a[:-k])
    if k == 0:
        out = np.zeros_like(X) + m
    else:
        out = eig_synthesis(Xc, W, m, k)
    return out, k, var


def raw_patch(X, tau):
    """Hard singular-value threshold (lowrank.py:170-184)."""
    U, s, Vt, m = svd_analysis(X)
    r = np.sum(s > tau)
    if r == 0:
        out = np.zeros_like(X) + m
    else:
        s[s < tau] = 0
        out = svd_synthesis(U, s, Vt, m, r)
    return out, r, np.nan


def nordic_threshold(full_patch_shape, n_frames, noise_std, n_iter=10):
    """NORDIC global threshold (lowrank.py:221-242).

    Mean over ``n_iter`` draws of the largest singular value of a
    ``prod(full_patch_shape) x n_frames`` standard normal matrix drawn from the
    global legacy ``np.random`` stream (the patch size includes the time
    length), times ``noise_std`` (an array is averaged; anything else that is
    not a float raises ``ValueError``).
    """
    smax = sum(
        max(svd(np.random.randn(np.prod(full_patch_shape), n_frames), compute_uv=False))
        for _ in range(n_iter)
    )
    smax /= n_iter
    if isinstance(noise_std, np.ndarray):
        noise_std = np.mean(noise_std)
    if not isinstance(noise_std, (float, np.floating)):
        raise ValueError(
            "For NORDIC the noise level must be either an"
            + " array or a float specifying the std in the volume.",
        )
    return noise_std * smax


def _x_of_y(y, beta):
    """Eq. (8) of Gavish & Donoho 2017, gated at the bulk edge (lowrank.py:250-255)."""
    t = y**2 - beta - 1
    return np.sqrt(0.5 * (t + np.sqrt((t**2) - (4 * beta)))) * (y >= (1 + np.sqrt(beta)))


def shrink_ope(y, beta=1):
    """Operator-norm shrinker (lowrank.py:258-260)."""
    return np.maximum(_x_of_y(y, beta), 0)


def shrink_nuc(y, beta=1):
    """Nuclear-norm shrinker (lowrank.py:263-269)."""
    x = _x_of_y(y, beta)
    return np.maximum(0, (x**4 - (np.sqrt(beta) * x * y) - beta)) / ((x**2) * y)


def shrink_fro(y, beta=1):
    """Frobenius shrinker, ungated (lowrank.py:272-280)."""
    return np.sqrt(np.maximum((((y**2) - beta - 1) ** 2 - 4 * beta), 0) / y)


SHRINKERS = {"fro": shrink_fro, "nuc": shrink_nuc, "ope": shrink_ope}


def optimal_patch(X, shrink, mpm, var_patch=None):
    """Optimal singular-value shrinkage (lowrank.py:379-407).

    ``var_patch`` is the a-priori variance patch (or None: noise level from the
    median singular value and the MP median).  Rank = last nonzero shrunk
    value + 1.
    """
    U, s, Vt, m = svd_analysis(X)
    if var_patch is not None:
        sigma = np.mean(np.sqrt(var_patch))
    else:
        sigma = np.median(s) / np.sqrt(X.shape[1] * mpm)
    c = np.sqrt(X.shape[1]) * sigma
    s2 = c * shrink(s / c, beta=X.shape[1] / X.shape[0])
    s2[np.isnan(s2)] = 0
    if np.any(s2):
        r = np.max(np.nonzero(s2)) + 1
        out = svd_synthesis(U, s2, Vt, m, r)
    else:
        r = 0
        out = np.zeros_like(X) + m
    return out, r, np.nan


# --------------------------------------------------------------------------
# Patch driver + recombination (space_time/base.py:177-270)
# --------------------------------------------------------------------------


def denoise(data, ps, po, patch_fn, recombination="weighted", mask=None,
            mask_threshold=50):
    """Selection -> C-order per-patch loop -> recombination -> normalisation.

    ``patch_fn(X, idx) -> (patch, rank, noise)``.  weighted: theta =
    1/(2+rank); average: theta = 1; center: one element per patch (the centre
    voxel at the centre frame), weights untouched.  Then ``out /= w`` for
    weighted/average and ``out[~mask] = 0``.  Returns ``(out, w, noise, rank)``
    with all-zero noise / rank maps (never written by the reference).
    """
    if recombination not in ("weighted", "average", "center"):
        raise ValueError("recombination must be one of 'weighted', 'average', 'center'")
    shape = data.shape
    src = StridedPatches(data, ps, po)
    out_arr = np.zeros(shape, dtype=data.dtype)
    w_arr = np.zeros(shape, dtype=np.float32)
    out = StridedPatches(out_arr, ps, po)
    w = StridedPatches(w_arr, ps, po)
    sel, full_mask = select_patches(mask, shape, ps, po, mask_threshold)
    centre = tuple(p // 2 for p in ps)
    n_space = np.prod(ps[:-1])
    for i in sel:
        X = src.get(i).reshape(n_space, -1)
        P, r, _ = patch_fn(X, i)
        P = np.reshape(P, ps)
        if recombination == "center":
            out.get(i)[centre] = P[centre]
        elif recombination == "weighted":
            theta = 1 / (2 + r)
            o = out.get(i)
            o += P * theta
            ww = w.get(i)
            ww += theta
        else:
            o = out.get(i)
            o += P
            ww = w.get(i)
            ww += 1
    if recombination in ("average", "weighted"):
        with np.errstate(invalid="ignore", divide="ignore"):
            out_arr /= w_arr
    out_arr[~full_mask] = 0
    return out_arr, w_arr, np.zeros(shape, np.float32), np.zeros(shape, np.int32)


def _prior_variance(noise_std, data_shape, ps, po):
    """Per-voxel prior variance broadcast over time (lowrank.py:94-101, 369-376)."""
    if isinstance(noise_std, (float, np.floating)):
        var = noise_std**2 * np.ones(data_shape[:-1])
    else:
        var = noise_std**2
    return StridedPatches(np.broadcast_to(var[..., None], data_shape), ps, po)


# --------------------------------------------------------------------------
# Functional entry points (denoise.py:15-281) with the reference signatures
# --------------------------------------------------------------------------


def mp_pca(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
           recombination="weighted", threshold_scale=1.0, progbar=None):
    ps, po = patch_params(patch_shape, patch_overlap, input_data.shape)
    return denoise(input_data, ps, po,
                   lambda X, i: mppca_patch(X, threshold_scale),
                   recombination, mask, mask_threshold)


def hybrid_pca(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
               noise_std=1.0, recombination="weighted", progbar=None):
    ps, po = patch_params(patch_shape, patch_overlap, input_data.shape)
    prior = _prior_variance(noise_std, input_data.shape, ps, po)
    return denoise(input_data, ps, po,
                   lambda X, i: hybrid_patch(X, np.mean(prior.get(i))),
                   recombination, mask, mask_threshold)


def raw_svt(input_data, patch_shape, patch_overlap, mask_threshold=50, mask=None,
            threshold=1.0, recombination="weighted", progbar=None):
    if recombination not in ("weighted", "average", "center"):
        raise ValueError("recombination must be one of 'weighted', 'average', 'center'")
    ps, po = patch_params(patch_shape, patch_overlap, input_data.shape)
    tau = threshold * 1.0
    return denoise(input_data, ps, po, lambda X, i: raw_patch(X, tau),
                   recombination, mask, mask_threshold)


def nordic(input_data, patch_shape, patch_overlap, mask_threshold=50, mask=None,
           noise_std=1.0, recombination="weighted", n_iter_threshold=10, progbar=None):
    if recombination not in ("weighted", "average", "center"):
        raise ValueError("recombination must be one of 'weighted', 'average', 'center'")
    ps, po = patch_params(patch_shape, patch_overlap, input_data.shape)
    tau = nordic_threshold(ps, input_data.shape[-1], noise_std, n_iter_threshold)
    return denoise(input_data, ps, po, lambda X, i: raw_patch(X, tau),
                   recombination, mask, mask_threshold)


def optimal_thresholding(input_data, patch_shape, patch_overlap, mask=None,
                         mask_threshold=50, loss="fro", noise_std=None,
                         recombination="weighted", eps_marshenko_pastur=1e-7,
                         progbar=None):
    if recombination not in ("weighted", "average", "center"):
        raise ValueError("recombination must be one of 'weighted', 'average', 'center'")
    shrink = SHRINKERS[loss]
    ps, po = patch_params(patch_shape, patch_overlap, input_data.shape)
    beta = (ps[-1] / np.prod(ps[:-1]) if len(ps) == len(input_data.shape)
            else input_data.shape[-1] / np.prod(ps))
    mpm = mp_median(beta=beta, eps=eps_marshenko_pastur)
    prior = None if noise_std is None else _prior_variance(noise_std, input_data.shape, ps, po)
    return denoise(
        input_data, ps, po,
        lambda X, i: optimal_patch(X, shrink, mpm, None if prior is None else prior.get(i)),
        recombination, mask, mask_threshold)


# --------------------------------------------------------------------------
# Sliding-window noise map (space_time/utils.py:207-242)
# --------------------------------------------------------------------------


def estimate_noise(noise_sequence, block_size=1):
    """Population std of each stride-1 block^d x T window, written at the
    window centre (utils.py:207-242, corners from get_patch_locs, utils.py:8-55).

    Window corners run 0..S-b in every spatial dimension; like the reference
    (whose corner list breaks with an IndexError), a spatial size <= b raises
    IndexError.  Voxels that are no window's centre are NaN here (the
    reference leaves them uninitialised).
    """
    vshape = noise_sequence.shape[:-1]
    b = int(block_size)
    corners = []
    for s in vshape:
        if s - b <= 0:
            raise IndexError("index -1 is out of bounds for axis 0 with size 0")
        corners.append(range(0, s - b + 1))
    out = np.full(vshape, np.nan)
    for c in itertools.product(*corners):
        win = tuple(slice(ci, ci + b) for ci in c)
        out[tuple(ci + b // 2 for ci in c)] = np.std(noise_sequence[win])
    return out


# --------------------------------------------------------------------------
# Adaptive trace-norm shrinkage (space_time/lowrank.py:410-636)
# --------------------------------------------------------------------------


def atn_shrink(s, gamma, tau):
    """s_i max(1 - (tau / s_i)^gamma, 0) (lowrank.py:459-461)."""
    return s * np.maximum(1 - (tau / s) ** gamma, 0)


def atn_cost(shape, method, s, gamma, sigma=None, tau=None):
    """SURE / GSURE cost of the adaptive trace-norm estimator (lowrank.py:410-442).

    ``shape`` = (rows, frames) of the Casorati matrix.  The free variable is
    log(tau) for "sure" / "gsure" and log(gamma - 1) for "qut".  The pair
    matrix D is held in float32 as in the reference (NumPy path).
    """
    nr, nc = shape
    if method == "qut":
        gamma = np.exp(gamma) + 1
    else:
        tau = np.exp(tau)
    sq = s**2
    k = len(s)
    dhat = atn_shrink(s, gamma, tau)
    num = s * dhat
    pair = np.zeros((k, k), dtype=np.float32)
    for i in range(k):
        den = sq[i] - sq
        den[i] = np.inf
        pair[i, :] = num[i] / den
    gradd = (1 + (gamma - 1) * (tau / s) ** gamma) * (s >= tau)
    div = np.sum(gradd + abs(nr - nc) * dhat / s) + 2 * np.sum(pair)
    rss = np.sum((dhat - s) ** 2)
    if method == "gsure":
        return rss / (1 - div / nr / nc) ** 2
    return (sigma**2) * ((-nr * nc) + (2 * div)) + rss


def adaptive_params(shape, s, stdest, method, gamma0, tau0, nbsim):
    """(gamma, tau) of lowrank.py:444-513: QUT = quantile of nbsim Monte-Carlo
    top singular values (global np.random) and, unless gamma0 is a float,
    gamma by BFGS on the cost; SURE / GSURE = BFGS on log(tau) from tau0
    (log median singular value) for each gamma candidate, lowest cost wins."""
    from scipy.optimize import minimize

    if method == "qut":
        maxd = np.ones(nbsim)
        for i in range(nbsim):
            maxd[i] = np.max(svd(np.random.randn(*shape) * stdest, compute_uv=False,
                                 overwrite_a=True))
        tau = np.quantile(maxd, 1 - 1 / np.sqrt(np.log(max(*shape))))
        if isinstance(gamma0, (float, np.floating)):
            return gamma0, tau
        res = minimize(lambda g: atn_cost(shape, "qut", s, g, stdest, tau), 0)
        return np.exp(res.x) + 1, tau
    if tau0 is None:
        tau0 = np.log(np.median(s))
    best = np.inf
    gamma = tau = None
    for g in gamma0:
        res = minimize(lambda x: atn_cost(shape, method, s, g, stdest, x), tau0)  # noqa: B023
        c = atn_cost(shape, method, s, g, stdest, res.x)
        if c < best:
            gamma, tau, best = g, np.exp(res.x), c
    return gamma, tau


def adaptive_patch(X, var_patch, method, gamma0, tau0, nbsim):
    """Adaptive shrinkage of one Casorati matrix (lowrank.py:609-636)."""
    stdest = np.sqrt(np.mean(var_patch))
    U, s, Vt, m = svd_analysis(X)
    gamma, tau = adaptive_params(X.shape, s, stdest, method, gamma0, tau0, nbsim)
    sh = atn_shrink(s, gamma, tau)
    if np.any(sh):
        r = np.max(np.nonzero(sh)) + 1
        out = svd_synthesis(U, sh, Vt, m, r)
    else:
        r = 0
        out = np.zeros_like(X) + m
    return out, r, np.nan


def adaptive_thresholding(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
                          recombination="weighted", method="SURE", nbsim=500, tau0=None,
                          gamma0=None, noise_std=1.0, progbar=None):
    """denoise.py:284-348 / AdaptiveDenoiser (lowrank.py:543-595)."""
    if recombination not in ("weighted", "average", "center"):
        raise ValueError("recombination must be one of 'weighted', 'average', 'center'")
    if method.lower() not in ("sure", "qut", "gsure"):
        raise ValueError(f"Unsupported method: '{method}'")
    ps, po = patch_params(patch_shape, patch_overlap, input_data.shape)
    prior = _prior_variance(noise_std, input_data.shape, ps, po)
    meth = method.lower()
    return denoise(input_data, ps, po,
                   lambda X, i: adaptive_patch(X, prior.get(i), meth, gamma0, tau0, nbsim),
                   recombination, mask, mask_threshold)
