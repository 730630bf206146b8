"""Automatic brain mask of an EPI series, for ``--mask auto`` / no ``--mask``.

The reference command line fits nilearn's ``NiftiMasker(mask_strategy="epi")``
(reference bindings/cli.py:408-413), i.e. nilearn's ``compute_epi_mask`` with
its defaults (lower_cutoff 0.2, upper_cutoff 0.85, connected, opening 2).
nilearn is not a dependency here; this is a restatement of that published
algorithm (Nichols' histogram-gap threshold):

1. mean image over time, non-finite values set to 0, smoothed with a 1 mm
   FWHM Gaussian (voxel sizes from the affine) because opening is on;
2. sort the mean image; between the 20% and 85% quantile positions find the
   largest gap between consecutive sorted values; threshold halfway across it;
3. erode twice, keep the largest 6-connected component, dilate four times and
   erode twice (the opening and closing of the post-processing).

Parity with nilearn is unpinned (nilearn is absent, and the reference's
test passes an explicit mask); the mask is written next to the output, as in
the reference, so it can be inspected.
"""

import logging

import numpy as np
from scipy import ndimage


def _smooth(mean, affine, fwhm=1.0):
    vox = np.sqrt((np.asarray(affine, np.float64)[:3, :3] ** 2).sum(axis=0)) if affine is not None \
        else np.ones(3)
    sigma = fwhm / np.sqrt(8.0 * np.log(2.0)) / vox[: mean.ndim]
    out = mean.astype(np.float64)
    for ax, s in enumerate(sigma):
        if s > 0:
            out = ndimage.gaussian_filter1d(out, s, axis=ax)
    return out


def compute_epi_mask(data, affine=None, lower_cutoff=0.2, upper_cutoff=0.85, opening=2):
    """Boolean mask over the spatial axes of ``data`` (the last axis is time)."""
    data = np.asarray(data)
    mean = np.abs(data).mean(axis=-1) if np.iscomplexobj(data) else data.mean(axis=-1)
    mean = np.where(np.isfinite(mean), mean, 0.0)
    if opening:
        mean = _smooth(mean, affine)
    v = np.sort(mean.ravel())
    lo = int(np.floor(lower_cutoff * len(v)))
    hi = min(int(np.floor(upper_cutoff * len(v))), len(v) - 1)
    delta = v[lo + 1:hi + 1] - v[lo:hi]
    if delta.size == 0:
        thr = v[lo]
    else:
        ia = int(delta.argmax())
        thr = 0.5 * (v[ia + lo] + v[ia + lo + 1])
    mask = mean >= thr
    if opening:
        mask = ndimage.binary_erosion(mask, iterations=opening)
    if mask.any():
        lab, n = ndimage.label(mask)
        if n > 1:
            sizes = np.bincount(lab.ravel())[1:]
            mask = lab == (1 + int(sizes.argmax()))
    else:
        logging.warning("Computed an empty mask")
    if opening:
        mask = ndimage.binary_dilation(mask, iterations=2 * opening)
        mask = ndimage.binary_erosion(mask, iterations=opening)
    return mask
