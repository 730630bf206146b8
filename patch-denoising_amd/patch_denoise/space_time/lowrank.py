"""Low-rank denoisers (mirrors space_time/lowrank.py) on the MI355X library.

Each class keeps the reference's constructor, ``denoise`` signature, stateful
``input_denoising_kwargs`` and exceptions; the per-patch work is the GPU rule
selected by ``_rule``.  Host code computes only the global scalars the
reference also computes once per call (NORDIC's Monte-Carlo threshold, the
Marchenko-Pastur median) and the prior variance map.
"""

from types import MappingProxyType

import numpy as np
from scipy.linalg import svd

from .. import _native as N
from .base import BaseSpaceTimeDenoiser, PatchedArray, prior_is_float32, spatial_variance_map
from .utils import marchenko_pastur_median


def _compares_in_f64(data_dtype, operand):
    """True when NumPy evaluates ``spectrum <op> operand`` in float64.

    The reference's spectra have the data's real precision; a Python float is
    a weak scalar (kept in float32 for float32 data), a NumPy float64 is not.
    """
    real = np.dtype(data_dtype).type(0).real.dtype
    return np.result_type(real, operand) == np.float64


class MPPCADenoiser(BaseSpaceTimeDenoiser):
    """Marchenko-Pastur PCA (reference: lowrank.py:27-63)."""

    _rule = N.MPPCA

    def __init__(self, patch_shape, patch_overlap, threshold_scale, **kwargs):
        super().__init__(patch_shape, patch_overlap, **kwargs)
        self.input_denoising_kwargs["threshold_scale"] = threshold_scale

    def _set_scale(self, prob, dtype, threshold_scale):
        prob.threshold_scale2 = float(threshold_scale**2)
        prob.scale_f64 = int(_compares_in_f64(dtype, threshold_scale**2))

    def _configure(self, prob, input_data, p_s):
        self._set_scale(prob, input_data.dtype, self.input_denoising_kwargs["threshold_scale"])
        return None

    def _patch_processing(self, patch, patch_idx=None, threshold_scale=1.0):
        """MP-PCA on one Casorati matrix, on the GPU (lowrank.py:42-63)."""

        def conf(prob, patches):
            self._set_scale(prob, patches.dtype, threshold_scale)

        out, r, v, _ = self._process_casorati(np.asarray(patch)[None], configure=conf)
        return out[0], int(r[0]), np.float32(v[0])


class HybridPCADenoiser(BaseSpaceTimeDenoiser):
    """Hybrid-PCA against a prior noise level (reference: lowrank.py:66-119)."""

    _rule = N.HYBRID

    def denoise(self, input_data, mask=None, mask_threshold=50, noise_std=1.0, progbar=None):
        input_data = np.asarray(input_data)
        p_s, p_o = self._get_patch_param(input_data.shape)
        var_map = spatial_variance_map(noise_std, input_data.shape)
        self.input_denoising_kwargs["var_apriori"] = PatchedArray(
            np.broadcast_to(var_map[..., None], input_data.shape), p_s, p_o
        )
        self._var_map = var_map
        self._var_f32 = prior_is_float32(noise_std)
        return super().denoise(input_data, mask, mask_threshold, progbar=progbar)

    def _configure(self, prob, input_data, p_s):
        prob.var_f32 = int(getattr(self, "_var_f32", False))
        return self._var_map

    def _patch_processing(self, patch, patch_idx=None, var_apriori=None):
        """Hybrid-PCA on one Casorati matrix, on the GPU (lowrank.py:104-119)."""
        patch = np.asarray(patch)
        varest = float(np.mean(var_apriori.get_patch(patch_idx)))
        rows = np.full((1, patch.shape[0]), varest)
        out, r, v, _ = self._process_casorati(patch[None], var_rows=rows,
                                              configure=lambda prob, p: None)
        return out[0], int(r[0]), np.float32(v[0])


class RawSVDDenoiser(BaseSpaceTimeDenoiser):
    """Hard singular-value thresholding (reference: lowrank.py:122-184)."""

    _rule = N.HARD

    def __init__(self, patch_shape, patch_overlap, threshold_value=1.0, recombination="weighted",
                 **kwargs):
        self._threshold_val = threshold_value
        super().__init__(patch_shape, patch_overlap, recombination, **kwargs)

    def denoise(self, input_data, mask=None, mask_threshold=50, threshold_scale=1.0,
                progbar=None):
        self._threshold = self._threshold_val * threshold_scale
        return super().denoise(input_data, mask, mask_threshold, progbar=progbar)

    def _set_tau(self, prob, dtype):
        prob.tau = float(self._threshold)
        prob.tau_f64 = int(_compares_in_f64(dtype, self._threshold))

    def _configure(self, prob, input_data, p_s):
        self._set_tau(prob, input_data.dtype)
        return None

    def _patch_processing(self, patch, patch_idx=None, **kwargs):
        """Hard SVT on one Casorati matrix, on the GPU (lowrank.py:170-184)."""
        out, r, _, _ = self._process_casorati(
            np.asarray(patch)[None], configure=lambda prob, p: self._set_tau(prob, p.dtype))
        return out[0], int(r[0]), np.nan


def nordic_threshold(n_rows, n_frames, n_iter):
    """Mean over ``n_iter`` draws of the top singular value of an
    ``n_rows x n_frames`` standard normal matrix from the global legacy
    ``np.random`` stream (reference: lowrank.py:221-232).

    Small draws use the same ``scipy.linalg.svd`` call as the reference; large
    ones (NORDIC's draws include the frames in the row count: 266,200 x 200
    for 11^3-voxel patches of 200 frames) go to the library's
    ``pdamd_top_singular_values`` (FP64 Gram, Householder, bisection: the
    same value to ~1e-15 relative).  The draws themselves stay on the host:
    they must be the global stream's own legacy Gaussian sequence.
    """
    total = 0.0
    for _ in range(n_iter):
        a = np.random.randn(n_rows, n_frames)
        if a.size <= 4_000_000:
            total = total + max(svd(a, compute_uv=False))
        else:
            total = total + float(N.top_singular_values(a))
    return np.float64(total) / n_iter


class NordicDenoiser(RawSVDDenoiser):
    """NORDIC (reference: lowrank.py:187-246)."""

    def denoise(self, input_data, mask=None, mask_threshold=50, noise_std=1.0,
                n_iter_threshold=10, progbar=None):
        input_data = np.asarray(input_data)
        patch_shape, _ = self._get_patch_param(input_data.shape)
        max_sval = nordic_threshold(int(np.prod(patch_shape)), input_data.shape[-1],
                                    n_iter_threshold)
        if isinstance(noise_std, np.ndarray):
            noise_std = np.mean(noise_std)
        if not isinstance(noise_std, (float, np.floating)):
            raise ValueError(
                "For NORDIC the noise level must be either an"
                + " array or a float specifying the std in the volume.",
            )
        self._threshold = noise_std * max_sval
        return BaseSpaceTimeDenoiser.denoise(self, input_data, mask,
                                             mask_threshold=mask_threshold, progbar=progbar)


class OptimalSVDDenoiser(BaseSpaceTimeDenoiser):
    """Optimal singular-value shrinkage (reference: lowrank.py:283-407)."""

    _rule = N.OPTIMAL
    _OPT_LOSS_SHRINK = MappingProxyType({"fro": "fro", "nuc": "nuc", "ope": "ope"})

    def __init__(self, patch_shape, patch_overlap, loss="fro", recombination="weighted",
                 **kwargs):
        super().__init__(patch_shape, patch_overlap, recombination=recombination, **kwargs)
        self.input_denoising_kwargs["shrink_func"] = OptimalSVDDenoiser._OPT_LOSS_SHRINK[loss]

    def denoise(self, input_data, mask=None, mask_threshold=50, noise_std=None,
                eps_marshenko_pastur=1e-7, progbar=None):
        input_data = np.asarray(input_data)
        p_s, p_o = self._get_patch_param(input_data.shape)
        self.input_denoising_kwargs["mp_median"] = marchenko_pastur_median(
            beta=(
                p_s[-1] / np.prod(p_s[:-1])
                if len(p_s) == len(input_data.shape)
                else input_data.shape[-1] / np.prod(p_s)
            ),
            eps=eps_marshenko_pastur,
        )
        self._var_map = None
        self._var_f32 = noise_std is not None and prior_is_float32(noise_std)
        if noise_std is None:
            self.input_denoising_kwargs["var_apriori"] = None
        else:
            self._var_map = spatial_variance_map(noise_std, input_data.shape)
            self.input_denoising_kwargs["var_apriori"] = PatchedArray(
                np.broadcast_to(self._var_map[..., None], input_data.shape), p_s, p_o
            )
        return super().denoise(input_data, mask, mask_threshold, progbar=progbar)

    def _set_rule(self, prob, mp_median, shrink_func):
        prob.loss = N.LOSS[shrink_func]
        prob.mp_median = float(mp_median)

    def _configure(self, prob, input_data, p_s):
        kw = self.input_denoising_kwargs
        self._set_rule(prob, kw["mp_median"], kw["shrink_func"])
        prob.var_f32 = int(getattr(self, "_var_f32", False))
        return self._var_map

    def _patch_processing(self, patch, patch_idx=None, shrink_func=None, mp_median=None,
                          var_apriori=None):
        """Optimal shrinkage of one Casorati matrix, on the GPU (lowrank.py:379-407)."""
        patch = np.asarray(patch)
        rows = None
        if var_apriori is not None:
            sigma = float(np.mean(np.sqrt(var_apriori.get_patch(patch_idx))))
            rows = np.full((1, patch.shape[0]), sigma * sigma)
        out, r, _, _ = self._process_casorati(
            patch[None], var_rows=rows,
            configure=lambda prob, p: self._set_rule(prob, mp_median, shrink_func))
        return out[0], int(r[0]), np.nan


def _qut_max_svals(n_rows, n_frames, stdest, nbsim):
    """Largest singular value of each of ``nbsim`` draws ``randn(n_rows,
    n_frames) * stdest`` from the global legacy ``np.random`` stream, in the
    reference's order (lowrank.py:446-456: one ``randn(*patch.shape)`` per
    draw; a single ``randn(nbsim, n_rows, n_frames)`` yields the same
    numbers).  The values come from the library's
    ``pdamd_top_singular_values`` over the whole stack (the float64 min-side
    Gram's top eigenvalue: the SVD's value to ~1e-15)."""
    a = np.random.randn(nbsim, n_rows, n_frames) * stdest
    return N.top_singular_values(a)


class AdaptiveDenoiser(BaseSpaceTimeDenoiser):
    """Adaptive trace-norm shrinkage, SURE / GSURE / QUT parameter selection
    (reference: lowrank.py:543-636, after Josse & Sardy).

    The per-patch parameter search runs on the GPU: SURE / GSURE minimise the
    reference's cost in log(tau) for each gamma candidate from log(tau0)
    (log median singular value by default) by downhill bracketing and
    golden-section search; QUT takes tau from the same Monte-Carlo draws as
    the reference (global ``np.random``, patch order) and, unless a float
    gamma0 is given, minimises the cost in log(gamma - 1) from 0.  The
    reference minimises with BFGS on finite-difference gradients of a cost
    whose pair matrix is rounded to float32, so parameters agree to the
    optimiser's precision, not bit for bit.
    """

    _rule = N.ADAPTIVE
    _SUPPORTED_METHOD = ["sure", "qut", "gsure"]

    def __init__(self, patch_shape, patch_overlap, method="SURE", recombination="weighted",
                 nbsim=500, **kwargs):
        super().__init__(patch_shape, patch_overlap, recombination, **kwargs)
        if method.lower() not in self._SUPPORTED_METHOD:
            raise ValueError(
                f"Unsupported method: '{method}', use any of {self._SUPPORTED_METHOD}"
            )
        self.input_denoising_kwargs["method"] = method.lower()
        self.input_denoising_kwargs["nbsim"] = nbsim

    def denoise(self, input_data, mask=None, mask_threshold=50, tau0=None, noise_std=None,
                gamma0=None, progbar=None):
        input_data = np.asarray(input_data)
        self.input_denoising_kwargs["gamma0"] = gamma0
        self.input_denoising_kwargs["tau0"] = tau0
        p_s, p_o = self._get_patch_param(input_data.shape)
        if isinstance(noise_std, (float, np.floating)):
            var_apriori = noise_std**2 * np.ones(input_data.shape[:-1])
        else:
            var_apriori = noise_std**2
        full = np.broadcast_to(var_apriori[..., None], input_data.shape)
        self.input_denoising_kwargs["var_apriori"] = PatchedArray(full, p_s, p_o)
        self._var_map = np.ascontiguousarray(full[..., 0], dtype=np.float64)
        self._var_f32 = np.asarray(var_apriori).dtype == np.float32
        self._mask_for_qut = (mask, mask_threshold)
        return super().denoise(input_data, mask, mask_threshold, progbar=progbar)

    # -- rule operands --------------------------------------------------------------
    def _gammas(self, method, gamma0):
        if method == "qut":
            return [float(gamma0)] if isinstance(gamma0, (float, np.floating)) else []
        return [float(g) for g in gamma0]  # None raises TypeError, as in the reference

    def _set_rule(self, prob, method, gamma0, tau0, qut_tau=None):
        import torch

        gam = self._gammas(method, gamma0)
        dev = torch.device("cuda", torch.cuda.current_device())
        self._gamma_t = torch.tensor(gam if gam else [0.0], dtype=torch.float64, device=dev)
        prob.adaptive_method = N.ADAPT[method]
        prob.n_gamma = len(gam)
        prob.gamma_dev = self._gamma_t.data_ptr()
        prob.tau0_log = float("nan") if tau0 is None else float(np.asarray(tau0).reshape(-1)[0])
        if qut_tau is not None:
            self._qut_t = torch.from_numpy(np.ascontiguousarray(qut_tau, np.float64)).to(dev)
            prob.qut_tau_dev = self._qut_t.data_ptr()

    def _qut_taus(self, prob, input_data, mask_arr, npatch):
        """QUT threshold of every selected patch, in patch order (lowrank.py:444-457)."""
        kw = self.input_denoising_kwargs
        idx = N.select_patches(prob, mask_arr)
        var_ap = kw["var_apriori"]
        taus = np.zeros(max(npatch, 1), np.float64)
        M = int(np.prod(prob.patch[:3]))
        T = int(prob.patch[3])
        q = 1 - 1 / np.sqrt(np.log(max(M, T)))
        for i in idx:
            stdest = np.sqrt(np.mean(var_ap.get_patch(int(i))))
            taus[i] = np.quantile(_qut_max_svals(M, T, stdest, kw["nbsim"]), q)
        return taus

    def _configure(self, prob, input_data, p_s):
        kw = self.input_denoising_kwargs
        qut = None
        self._set_rule(prob, kw["method"], kw["gamma0"], kw["tau0"])  # valid for the selection
        if kw["method"] == "qut":
            npatch = 1
            for i in range(4):
                S, p, st = int(prob.shape[i]), int(prob.patch[i]), int(prob.step[i])
                npatch *= ((S - p) // st + 1) if (p < S and st > 0) else 1
            mask, _ = self._mask_for_qut
            mask_arr = None if mask is None else np.asarray(mask)
            qut = self._qut_taus(prob, input_data, mask_arr, npatch)
        self._set_rule(prob, kw["method"], kw["gamma0"], kw["tau0"], qut)
        prob.var_f32 = int(getattr(self, "_var_f32", False))
        return self._var_map

    def _patch_processing(self, patch, patch_idx=None, gamma0=None, tau0=None, var_apriori=None,
                          method=None, nbsim=None):
        """Adaptive shrinkage of one Casorati matrix, on the GPU (lowrank.py:597-636)."""
        patch = np.asarray(patch)
        stdest = np.sqrt(np.mean(var_apriori.get_patch(patch_idx)))
        rows = np.full((1, patch.shape[0]), stdest * stdest)
        qut = None
        if method == "qut":
            M, T = patch.shape
            qut = np.array([np.quantile(_qut_max_svals(M, T, stdest, nbsim),
                                        1 - 1 / np.sqrt(np.log(max(M, T))))])
        out, r, _, _ = self._process_casorati(
            patch[None], var_rows=rows,
            configure=lambda prob, p: self._set_rule(prob, method, gamma0, tau0, qut))
        return out[0], int(r[0]), np.nan
