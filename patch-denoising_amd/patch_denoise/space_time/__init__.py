"""Space-time low-rank denoisers running on MI355X (mirrors space_time/__init__.py)."""

from .base import BaseSpaceTimeDenoiser, PatchedArray
from .lowrank import (
    AdaptiveDenoiser,
    HybridPCADenoiser,
    MPPCADenoiser,
    NordicDenoiser,
    OptimalSVDDenoiser,
    RawSVDDenoiser,
)

__all__ = [
    "BaseSpaceTimeDenoiser",
    "PatchedArray",
    "MPPCADenoiser",
    "HybridPCADenoiser",
    "NordicDenoiser",
    "OptimalSVDDenoiser",
    "RawSVDDenoiser",
    "AdaptiveDenoiser",
]
