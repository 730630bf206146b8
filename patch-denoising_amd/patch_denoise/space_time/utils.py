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
