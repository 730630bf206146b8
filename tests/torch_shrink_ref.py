"""Autograd reference for the differentiable optimal-shrinkage path (tests only).

Float64 PyTorch on the CPU: per patch a centred ``torch.linalg.svd``, the
optimal shrinker of space_time/lowrank.py:250-280 with the noise level of
lowrank.py:387-395 scaled by the patch's ``lambda``, and the weighted /
average recombination of space_time/base.py:229-262.  Gradients come from
torch's own SVD backward, independent of the hand-written adjoint kernels.
Spatial stride = patch - overlap along each axis; one temporal patch (the
whole series), as in the optimal-thresholding workloads.
"""

import itertools
import math

import torch


def _shrink(loss, y, beta):
    t = y * y - beta - 1.0
    if loss == "fro":
        Z = t * t - 4.0 * beta
        ok = Z > 0
        return torch.where(ok, torch.sqrt(torch.where(ok, Z, torch.ones_like(Z)) / y),
                           torch.zeros_like(y))
    D2 = t * t - 4.0 * beta
    ok = (y >= 1.0 + math.sqrt(beta)) & (D2 > 0)
    D2s = torch.where(ok, D2, torch.ones_like(D2))
    x = torch.where(ok, torch.sqrt(0.5 * (torch.where(ok, t, torch.ones_like(t)) + torch.sqrt(D2s))),
                    torch.zeros_like(y))
    if loss == "ope":
        return x
    xs = torch.where(ok, x, torch.ones_like(x))
    N = xs ** 4 - math.sqrt(beta) * xs * y - beta
    ok2 = ok & (N > 0)
    return torch.where(ok2, torch.where(ok2, N, torch.ones_like(N)) / (xs * xs * y),
                       torch.zeros_like(y))


def optimal_shrinkage_ref(x, lam, ps, po, mpm, loss="fro", recombination="weighted",
                          prior_std=None):
    """x: (X, Y, Z, T) float64 tensor; lam: patch-grid tensor; ps, po: spatial
    patch shape and overlap (3-tuples).  prior_std: None (median rule) or an
    (X, Y, Z) tensor of noise standard deviations.  Returns the denoised volume
    (NaN where no patch wrote, as the reference)."""
    X, Y, Zd, T = x.shape
    st = [p - o for p, o in zip(ps, po)]
    grid = [(S - p) // s + 1 for S, p, s in zip((X, Y, Zd), ps, st)]
    M = ps[0] * ps[1] * ps[2]
    beta = T / M
    out = torch.zeros_like(x)
    w = torch.zeros(x.shape, dtype=torch.float64)
    lam = lam.reshape(grid)
    for g in itertools.product(*(range(v) for v in grid)):
        o = [gi * si for gi, si in zip(g, st)]
        sl = tuple(slice(oi, oi + p) for oi, p in zip(o, ps))
        P = x[sl].reshape(M, T)
        m = P.mean(0)
        Xc = P - m
        U, s, Vh = torch.linalg.svd(Xc, full_matrices=False)
        if prior_std is None:
            ss = torch.sort(s).values
        n = s.shape[0]
        if prior_std is None:
            med = ss[(n - 1) // 2] if n % 2 else 0.5 * (ss[n // 2 - 1] + ss[n // 2])
            sigma = med / math.sqrt(T * mpm)
        else:
            sigma = prior_std[sl].mean()
        c = lam[g] * math.sqrt(T) * sigma
        gsv = c * _shrink(loss, s / c, beta)
        nz = torch.nonzero(gsv.detach() != 0)
        r = int(nz.max()) + 1 if nz.numel() else 0
        Pd = (U * gsv) @ Vh + m
        th = 1.0 / (2 + r) if recombination == "weighted" else 1.0
        out[sl] = out[sl] + th * Pd.reshape(ps + (T,))
        w[sl] = w[sl] + th
    return out / w
