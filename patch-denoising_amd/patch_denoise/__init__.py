"""patch_denoise on MI355X: space-time low-rank denoising with HIP kernels.

Drop-in for the reference package's low-rank entry points (reference
``__init__.py:3-33``, without ``adaptive_thresholding``): the functional API
and the denoiser classes keep their names and signatures; the per-patch work
runs in ``_lib/libpdamd.so`` on the GPU.
"""

from .denoise import hybrid_pca, mp_pca, nordic, optimal_thresholding, raw_svt
from .space_time.lowrank import (
    HybridPCADenoiser,
    MPPCADenoiser,
    NordicDenoiser,
    OptimalSVDDenoiser,
    RawSVDDenoiser,
)

__all__ = [
    "HybridPCADenoiser",
    "MPPCADenoiser",
    "NordicDenoiser",
    "OptimalSVDDenoiser",
    "RawSVDDenoiser",
    "mp_pca",
    "hybrid_pca",
    "optimal_thresholding",
    "raw_svt",
    "nordic",
]

__version__ = "0.1.0+mi355x"
