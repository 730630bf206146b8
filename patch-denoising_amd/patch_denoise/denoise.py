"""Functional entry points (mirrors the reference's denoise.py:15-281).

Same names, positional order (including ``raw_svt`` / ``nordic`` taking
``mask_threshold`` before ``mask``), defaults and 4-tuple return as the
reference.  ``precision`` is an extra keyword-only option ("auto", "fp32",
"fp64") for the GPU eigen stage.
"""

from .space_time.lowrank import (
    AdaptiveDenoiser,
    HybridPCADenoiser,
    MPPCADenoiser,
    NordicDenoiser,
    OptimalSVDDenoiser,
    RawSVDDenoiser,
)


def mp_pca(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
           recombination="weighted", threshold_scale=1.0, progbar=None, *, precision="auto"):
    """Marchenko-Pastur PCA denoising (reference denoise.py:15-62).

    As in the reference, ``progbar`` is accepted but not forwarded.
    """
    denoiser = MPPCADenoiser(patch_shape, patch_overlap, recombination=recombination,
                             threshold_scale=threshold_scale, precision=precision)
    return denoiser.denoise(input_data, mask=mask, mask_threshold=mask_threshold)


def hybrid_pca(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
               noise_std=1.0, recombination="weighted", progbar=None, *, precision="auto"):
    """Hybrid-PCA denoising (reference denoise.py:66-112)."""
    denoiser = HybridPCADenoiser(patch_shape, patch_overlap, recombination=recombination,
                                 precision=precision)
    return denoiser.denoise(input_data, mask=mask, mask_threshold=mask_threshold,
                            noise_std=noise_std, progbar=progbar)


def raw_svt(input_data, patch_shape, patch_overlap, mask_threshold=50, mask=None,
            threshold=1.0, recombination="weighted", progbar=None, *, precision="auto"):
    """Hard singular-value thresholding (reference denoise.py:116-163)."""
    denoiser = RawSVDDenoiser(patch_shape, patch_overlap, recombination=recombination,
                              threshold_value=threshold, precision=precision)
    return denoiser.denoise(input_data, mask=mask, mask_threshold=mask_threshold,
                            threshold_scale=1.0, progbar=progbar)


def nordic(input_data, patch_shape, patch_overlap, mask_threshold=50, mask=None,
           noise_std=1.0, recombination="weighted", n_iter_threshold=10, progbar=None, *,
           precision="auto"):
    """NORDIC denoising (reference denoise.py:167-221)."""
    denoiser = NordicDenoiser(patch_shape, patch_overlap, recombination=recombination,
                              precision=precision)
    return denoiser.denoise(input_data, mask=mask, mask_threshold=mask_threshold,
                            noise_std=noise_std, n_iter_threshold=n_iter_threshold,
                            progbar=progbar)


def optimal_thresholding(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
                         loss="fro", noise_std=None, recombination="weighted",
                         eps_marshenko_pastur=1e-7, progbar=None, *, precision="auto"):
    """Optimal singular-value shrinkage (reference denoise.py:225-281)."""
    denoiser = OptimalSVDDenoiser(patch_shape, patch_overlap, recombination=recombination,
                                  loss=loss, precision=precision)
    return denoiser.denoise(input_data, mask=mask, noise_std=noise_std,
                            mask_threshold=mask_threshold,
                            eps_marshenko_pastur=eps_marshenko_pastur, progbar=progbar)


def adaptive_thresholding(input_data, patch_shape, patch_overlap, mask=None, mask_threshold=50,
                          recombination="weighted", method="SURE", nbsim=500, tau0=None,
                          gamma0=None, noise_std=1.0, progbar=None, *, precision="auto"):
    """Adaptive trace-norm shrinkage, SURE / GSURE / QUT (reference denoise.py:284-348)."""
    denoiser = AdaptiveDenoiser(patch_shape, patch_overlap, recombination=recombination,
                                method=method, nbsim=nbsim, precision=precision)
    return denoiser.denoise(input_data, mask, mask_threshold, tau0, noise_std, gamma0,
                            progbar=progbar)
