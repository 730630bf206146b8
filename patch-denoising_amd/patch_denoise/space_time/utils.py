"""Host-side scalar helpers of the space-time path (mirrors space_time/utils.py).

Only the global scalar the GPU rules consume lives here: the Marchenko-Pastur
median used by optimal shrinkage (reference: utils.py:149-204).  The per-patch
analysis / synthesis of utils.py:58-146 runs inside the GPU kernels.
"""

import numpy as np
from scipy.integrate import quad


def marchenko_pastur_median(beta, eps=1e-7):
    """Median of the Marchenko-Pastur law of aspect ratio ``beta`` in [0, 1].

    Same bracketing as the reference: each round evaluates the CDF (adaptive
    quadrature of the MP density from the lower edge) at 5 equispaced points
    of the bracket and keeps the points straddling 1/2; it stops when the
    bracket is narrower than ``eps``, stops moving, or after 20 rounds, and
    returns the bracket midpoint.  ``ValueError`` outside [0, 1].
    """
    if not (0 <= beta <= 1):
        raise ValueError("Aspect Ratio should be between 0 and 1")
    upper = (1 + np.sqrt(beta)) ** 2
    lower = (1 - np.sqrt(beta)) ** 2

    def density(x):
        if upper >= x >= lower:
            return np.sqrt((upper - x) * (x - lower)) / (2 * np.pi * x * beta)
        return 0

    lo, hi = lower, upper
    for _ in range(20):
        if not (hi - lo) > eps:
            break
        pts = np.linspace(lo, hi, 5)
        cdf = np.array([quad(density, lower, x)[0] for x in pts])
        moved = False
        if np.any(cdf < 0.5):
            lo = np.max(pts[cdf < 0.5])
            moved = True
        if np.any(cdf > 0.5):
            hi = np.min(pts[cdf > 0.5])
            moved = True
        if not moved:
            break
    return (lo + hi) / 2


def estimate_noise(noise_sequence, block_size=1):
    """Estimate a noise map from a noise-only sequence (reference:
    space_time/utils.py:207-242), on the GPU.

    Each voxel that is the centre of a stride-1 window of ``block_size`` voxels
    per spatial axis gets the standard deviation (ddof 0) of the window's
    voxels over all frames.  Other voxels are NaN (the reference leaves them
    uninitialised).  A spatial size <= ``block_size`` raises IndexError, as the
    reference's patch-corner list does.

    Parameters
    ----------
    noise_sequence : np.ndarray of shape (X, Y, Z, T) (1 to 3 spatial axes)
    block_size : int

    Returns
    -------
    np.ndarray of the spatial shape, float64
    """
    from .. import _native as N

    noise_sequence = np.asarray(noise_sequence)
    for s in noise_sequence.shape[:-1]:
        if s - block_size <= 0:
            raise IndexError("index -1 is out of bounds for axis 0 with size 0")
    return N.estimate_noise(noise_sequence, block_size).cpu().numpy()
