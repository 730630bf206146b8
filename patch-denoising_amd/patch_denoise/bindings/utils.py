"""Denoiser registry, parameter strings and file I/O for the command line.

Mirrors reference bindings/utils.py: ``DENOISER_MAP`` (:20-41),
``DenoiseParameters`` (:46-113), ``load_as_array`` / ``save_array``
(:116-144) and ``load_complex_nifti`` (:147-175).  NIfTI files go through
``patch_denoise.bindings.nifti`` (the reference uses nibabel, which is not a
dependency here).
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ..denoise import (
    adaptive_thresholding,
    hybrid_pca,
    mp_pca,
    nordic,
    optimal_thresholding,
    raw_svt,
)
from . import nifti


def _optimal(loss):
    def run(*args, **kwargs):
        return optimal_thresholding(*args, loss=loss, **kwargs)

    run.__name__ = f"optimal_thresholding_{loss}"
    return run


def _adaptive_qut(*args, **kwargs):
    return adaptive_thresholding(*args, method="qut", **kwargs)


# command-line method name -> functional API call
DENOISER_MAP = {
    None: None,
    "mp-pca": mp_pca,
    "hybrid-pca": hybrid_pca,
    "raw": raw_svt,
    "optimal-fro": _optimal("fro"),
    "optimal-fro-noise": _optimal("fro"),
    "optimal-nuc": _optimal("nuc"),
    "optimal-ope": _optimal("ope"),
    "nordic": nordic,
    "adaptive-qut": _adaptive_qut,
}

# methods whose call needs a noise map (reference cli.py:487-495)
NOISE_MAP_METHODS = ("nordic", "hybrid-pca", "adaptive-qut", "optimal-fro-noise")


@dataclass
class DenoiseParameters:
    """A denoising setup, printable as ``<method>_<patch>_<overlap>_<recombination>_<mask threshold>``."""

    method: str | None = None
    patch_shape: int | tuple[int, ...] = 11
    patch_overlap: int | tuple[int, ...] = 0
    recombination: str = "weighted"
    mask_threshold: int = 10

    @property
    def pretty_name(self):
        """Method followed by every set (truthy) field, joined by ``_``; ``noisy`` without a method."""
        if not self.method:
            return "noisy"
        parts = [self.method]
        for f in ("patch_shape", "patch_overlap", "recombination", "mask_threshold"):
            v = getattr(self, f)
            if v:
                parts.append(str(v))
        return "_".join(parts)

    @property
    def pretty_par(self):
        """``<patch>_<overlap><first letter of recombination>``."""
        return f"{self.patch_shape}_{self.patch_overlap}{self.recombination[0]}"

    @classmethod
    def get_str(cls, **kwargs):
        return cls(**kwargs).pretty_name

    @classmethod
    def from_str(cls, config_str):
        """Parse a ``pretty_name`` string; missing trailing fields keep their defaults."""
        if "noisy" in config_str:
            return cls(method=None, patch_shape=None, patch_overlap=None,
                       recombination=None, mask_threshold=None)
        fields = config_str.split("_")
        d = cls()
        conv = (("method", str), ("patch_shape", int), ("patch_overlap", int),
                ("recombination", str), ("mask_threshold", int))
        for (name, typ), tok in zip(conv, fields):
            setattr(d, name, typ(tok))
        return d

    def __str__(self):
        return self.pretty_name


def _suffixes(path):
    return Path(path).suffixes


def load_as_array(input):
    """Load ``.npy`` or NIfTI-1 (``.nii``, ``.nii.gz``): (array, affine or None)."""
    if input is None:
        return None, None
    input = Path(input)
    if input.suffix == ".npy":
        return np.load(input), None
    if ".nii" in input.suffixes:
        return nifti.load(input, np.float32)
    raise ValueError("Unsupported file format. use numpy or nifti formats.")


def save_array(data, affine, filename):
    """Save to NIfTI-1 (with ``affine``, identity if None) or ``.npy``; None is a no-op."""
    if filename is None:
        return None
    filename = Path(filename)
    if ".nii" in filename.suffixes:
        if affine is None:
            affine = np.eye(len(data.shape))
        nifti.save(data, affine, filename)
    elif filename.suffix == ".npy":
        np.save(filename, data)
    return filename


def load_complex_nifti(mag_file, phase_file, filename=None):
    """Magnitude and phase (radians) files -> ``mag * exp(1j * phase)`` and the magnitude's affine."""
    mag, mag_affine = load_as_array(mag_file)
    phase, phase_affine = load_as_array(phase_file)
    if (mag_affine is None) != (phase_affine is None) or (
            mag_affine is not None and not np.allclose(mag_affine, phase_affine)):
        logging.warning("Affine matrices for magnitude and phase are not the same")
    logging.info("Phase data range is [%.2f %.2f]", np.min(phase), np.max(phase))
    logging.info("Mag data range is [%.2f %.2f]", np.min(mag), np.max(mag))
    img = mag * np.exp(1j * phase)
    if filename is not None:
        np.save(filename, img)
    return img, mag_affine
