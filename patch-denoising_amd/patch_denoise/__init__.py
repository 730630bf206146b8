"""patch_denoise on MI355X: space-time low-rank denoising with HIP kernels.

Drop-in for the reference package's low-rank entry points (reference
``__init__.py:3-33``): the functional API
and the denoiser classes keep their names and signatures; the per-patch work
runs in ``_lib/libpdamd.so`` on the GPU.
"""

from .denoise import (
    adaptive_thresholding,
    hybrid_pca,
    mp_pca,
    nordic,
    optimal_thresholding,
    raw_svt,
)
from .space_time.lowrank import (
    AdaptiveDenoiser,
    HybridPCADenoiser,
    MPPCADenoiser,
    NordicDenoiser,
    OptimalSVDDenoiser,
    RawSVDDenoiser,
)

__all__ = [
    "AdaptiveDenoiser",
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
    "adaptive_thresholding",
]

__version__ = "0.1.0+mi355x"
