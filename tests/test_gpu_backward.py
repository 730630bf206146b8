"""GPU tests of the differentiable optimal-shrinkage path.

The hand-written adjoint (pdamd_optimal_backward) is compared with
* the float64 autograd reference (torch.linalg.svd backward, tests/torch_shrink_ref.py),
* central finite differences (torch.autograd.gradcheck, eps = 1e-6, FP64),
and its forward at lambda = 1 with the oracle.  Sizes: the 8x8x8x20 (patch 3)
sub-volume of the rank-6 workload, and windows of the full 64x64x32x120 run.
"""

import numpy as np
import pytest
import torch

from tests.torch_shrink_ref import optimal_shrinkage_ref

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _sub_volume(shape=(8, 8, 8), T=20, seed=5):
    from patch_denoise import workloads

    noisy, clean = workloads.config5(shape=shape, T=T, seed=seed)
    return noisy.astype(np.float64), clean.astype(np.float64)


def _mpm(T, M):
    from oracle import cpu_ref as O

    return O.mp_median(T / M)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forward_lambda_one_matches_oracle(dtype):
    from oracle import cpu_ref as O
    from patch_denoise.differentiable import optimal_shrinkage

    noisy, _ = _sub_volume()
    data = noisy.astype(dtype)
    x = torch.from_numpy(data).cuda()
    lam = torch.ones(6 * 6 * 6, dtype=torch.float64, device="cuda")
    print(f"[bwd] forward lambda=1 {np.dtype(dtype).name}", flush=True)
    out = optimal_shrinkage(x, lam, 3, 2).cpu().numpy()
    mpm = _mpm(20, 27)
    ref, _, _, _ = O.denoise(data, (3, 3, 3, 20), (2, 2, 2, 20),
                             lambda X, i: O.optimal_patch(X, O.shrink_fro, mpm))
    e = rel_l2(out, ref)
    print(f"[bwd] rel-L2 vs oracle {e:.2e}", flush=True)
    # float64: the output is divided by float32 weights accumulated in another
    # order than the reference's C-order loop (~1e-7 relative)
    assert e < (5e-7 if dtype == np.float64 else 1e-5)


@pytest.mark.parametrize("loss,recomb,prior", [("fro", "weighted", False), ("nuc", "average", False),
                                               ("ope", "weighted", True), ("fro", "average", True)])
def test_backward_matches_svd_autograd(loss, recomb, prior):
    from patch_denoise.differentiable import optimal_shrinkage

    noisy, _ = _sub_volume()
    rng = np.random.default_rng(11)
    lam0 = 1.0 + 0.1 * rng.standard_normal(216)
    noise_std = (40.0 * (1.0 + 0.2 * rng.random((8, 8, 8)))) if prior else None
    # reference: float64 autograd through torch.linalg.svd on the CPU
    xr = torch.from_numpy(noisy).requires_grad_(True)
    lr = torch.from_numpy(lam0).requires_grad_(True)
    pstd = torch.from_numpy(noise_std) if prior else None
    outr = optimal_shrinkage_ref(xr, lr, (3, 3, 3), (2, 2, 2), _mpm(20, 27), loss, recomb, pstd)
    # a linear loss: the float32 recombination weights (rounded and summed in
    # another order than the reference's, ~1e-7 relative) then bound the
    # difference, instead of being amplified by a residual against the target
    gout = torch.from_numpy(np.random.default_rng(12).standard_normal(noisy.shape))
    lossr = (outr * gout).sum()
    lossr.backward()
    # GPU path
    x = torch.from_numpy(noisy).cuda().requires_grad_(True)
    lam = torch.from_numpy(lam0).cuda().requires_grad_(True)
    print(f"[bwd] {loss} {recomb} prior={prior}", flush=True)
    out = optimal_shrinkage(x, lam, 3, 2, loss=loss, recombination=recomb,
                            noise_std=noise_std)
    lg = (out * gout.cuda()).sum()
    lg.backward()
    ef = rel_l2(out.detach().cpu(), outr.detach())
    ex = rel_l2(x.grad.cpu(), xr.grad)
    el = rel_l2(lam.grad.cpu(), lr.grad)
    print(f"[bwd] forward {ef:.2e} dL/dx {ex:.2e} dL/dlambda {el:.2e}", flush=True)
    assert ef < 5e-7 and ex < 1e-6 and el < 1e-6


def _nuc_zero(beta):
    """y where the nuclear-norm shrinker leaves zero (N(x(y), y) = 0)."""
    lo, hi = 1.0 + np.sqrt(beta), 1.0 + np.sqrt(beta) + 10.0
    for _ in range(200):
        y = 0.5 * (lo + hi)
        t = y * y - beta - 1.0
        x = np.sqrt(0.5 * (t + np.sqrt(t * t - 4.0 * beta)))
        if x ** 4 - np.sqrt(beta) * x * y - beta > 0:
            hi = y
        else:
            lo = y
    return 0.5 * (lo + hi)


def _smooth_lambda(data, ps, loss, margin=0.02):
    """Per-patch lambda in [0.8, 1.25] keeping every y = s / (lambda c0) at
    least ``margin`` (relative) away from the rule's non-smooth points (bulk
    edges for fro / ope, the zero crossing for nuc), so that finite
    differences see a smooth function with a constant rank."""
    X, Y, Z, T = data.shape
    M = ps ** 3
    beta = T / M
    mpm = _mpm(T, M)
    pts = {"fro": [1 - np.sqrt(beta), 1 + np.sqrt(beta)], "ope": [1 + np.sqrt(beta)],
           "nuc": [_nuc_zero(beta)]}[loss]
    cands = np.linspace(0.8, 1.25, 181)
    lam, worst = [], np.inf
    for i in range(X - ps + 1):
        for j in range(Y - ps + 1):
            for k in range(Z - ps + 1):
                P = data[i:i + ps, j:j + ps, k:k + ps].reshape(M, T)
                sv = np.linalg.svd(P - P.mean(0), compute_uv=False)
                c0 = np.sqrt(T) * np.median(sv) / np.sqrt(T * mpm)
                best, bl = -1.0, 1.0
                for lc in cands:
                    y = sv / (lc * c0)
                    m = min(np.min(np.abs(y - q)) / q for q in pts)
                    if m > best:
                        best, bl = m, lc
                lam.append(bl)
                worst = min(worst, best)
    assert worst > margin, f"no lambda keeps a {margin} margin (best {worst})"
    return np.array(lam)


def _directional_errors(f, x, lam, eps, ndir, seed):
    """Fast-mode finite-difference check written out: v^T J u from the
    hand-written backward against central differences along random u."""
    rng = np.random.default_rng(seed)
    y = f(x, lam)
    errs = []
    for k in range(ndir):
        v = torch.from_numpy(rng.standard_normal(tuple(y.shape))).to(y)
        ux = torch.from_numpy(rng.standard_normal(tuple(x.shape))).to(x)
        ul = torch.from_numpy(rng.standard_normal(tuple(lam.shape))).to(lam) * 0.01
        gx, gl = torch.autograd.grad(y, (x, lam), v, retain_graph=True)
        an = float((gx * ux).sum() + (gl * ul).sum())
        with torch.no_grad():
            fp = float((f(x + eps * ux, lam + eps * ul) * v).sum())
            fm = float((f(x - eps * ux, lam - eps * ul) * v).sum())
        fd = (fp - fm) / (2 * eps)
        errs.append(abs(fd - an) / max(abs(an), 1e-30))
        print(f"[bwd] direction {k}: finite difference {fd:.10e} analytic {an:.10e}", flush=True)
    return errs


@pytest.mark.parametrize("loss,recomb", [("fro", "weighted"), ("nuc", "weighted"),
                                         ("ope", "average")])
def test_finite_differences_fp64_8x8x8x20(loss, recomb):
    """FP64, eps = 1e-6, 8x8x8x20 sub-volume, patch 3 (216 patches)."""
    from patch_denoise.differentiable import optimal_shrinkage

    noisy, _ = _sub_volume()
    x = torch.from_numpy(noisy).cuda().requires_grad_(True)
    lam = torch.from_numpy(_smooth_lambda(noisy, 3, loss)).cuda().requires_grad_(True)
    print(f"[bwd] finite differences {loss} {recomb}", flush=True)
    errs = _directional_errors(
        lambda a, b: optimal_shrinkage(a, b, 3, 2, loss=loss, recombination=recomb), x, lam,
        1e-6, 4, 3)
    # The forward's own rounding sets the floor: Gram eigenvalues carry
    # ulp * |G| absolute error, which the (ungated) Frobenius rule amplifies
    # for the smallest singular values; 1e-6 steps on values near 500.
    assert max(errs) < 1e-4


def test_gradcheck_full_jacobian_fp64():
    """torch.autograd.gradcheck (full Jacobians, eps = 1e-6) on a 5x5x5x12
    volume, patch 3: 27 patches, 1500 + 27 inputs."""
    from patch_denoise.differentiable import optimal_shrinkage

    noisy, _ = _sub_volume(shape=(5, 5, 5), T=12, seed=7)
    x = torch.from_numpy(noisy).cuda().requires_grad_(True)
    lam = torch.from_numpy(_smooth_lambda(noisy, 3, "nuc")).cuda().requires_grad_(True)
    print("[bwd] gradcheck full Jacobian (nuc, weighted)", flush=True)
    ok = torch.autograd.gradcheck(lambda a, b: optimal_shrinkage(a, b, 3, 2, loss="nuc"),
                                  (x, lam), eps=1e-6, atol=1e-5, rtol=1e-4,
                                  nondet_tol=1e-10,  # atomic scatter order varies
                                  raise_exception=False)
    print(f"[bwd] gradcheck: {ok}", flush=True)
    assert ok


def test_rank6_64x64x32x120_window_gradients():
    """Full workload in float32 (FP64 eigen stage): fwd + bwd, gradients on a
    window checked against the autograd reference on the sub-volume that holds
    every patch touching the window."""
    from patch_denoise import workloads
    from patch_denoise.differentiable import optimal_shrinkage

    noisy, clean = workloads.config5()
    x = torch.from_numpy(noisy).cuda().requires_grad_(True)
    grid = (60, 60, 28)
    lam = torch.ones(grid, dtype=torch.float32, device="cuda", requires_grad=True)
    lo, hi = (30, 30, 14), (33, 33, 17)
    wmask = torch.zeros(noisy.shape, dtype=torch.float32)
    wmask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1.0
    tgt = torch.from_numpy(clean)
    print("[bwd] 64x64x32x120 forward + backward", flush=True)
    out = optimal_shrinkage(x, lam, 5, 4)
    loss = (((out - tgt.cuda()) ** 2) * wmask.cuda()).sum()
    loss.backward()
    assert torch.isfinite(x.grad).all() and torch.isfinite(lam.grad).all()
    # reference on the sub-volume lo-4 .. hi+4 (patches touching the window)
    a = [l - 4 for l in lo]
    b = [h + 4 for h in hi]
    sub = torch.from_numpy(noisy[a[0]:b[0], a[1]:b[1], a[2]:b[2]].astype(np.float64))
    sub.requires_grad_(True)
    sg = [bb - aa - 4 for aa, bb in zip(a, b)]
    lr = torch.ones(sg, dtype=torch.float64, requires_grad=True)
    outr = optimal_shrinkage_ref(sub, lr, (5, 5, 5), (4, 4, 4), _mpm(120, 125), "fro")
    wl = wmask[a[0]:b[0], a[1]:b[1], a[2]:b[2]].double()
    tl = tgt[a[0]:b[0], a[1]:b[1], a[2]:b[2]].double()
    lref = (((outr - tl) ** 2) * wl).sum()
    lref.backward()
    # every patch touching the window lies in the sub-volume, so the gradient
    # is exact on the whole sub-volume (and zero outside it)
    gx = x.grad[a[0]:b[0], a[1]:b[1], a[2]:b[2]].cpu().double()
    gr = sub.grad
    gl = lam.grad[a[0]:a[0] + sg[0], a[1]:a[1] + sg[1], a[2]:a[2] + sg[2]].cpu().double()
    ex, el = rel_l2(gx, gr), rel_l2(gl, lr.grad)
    print(f"[bwd] window: dL/dx {ex:.2e} dL/dlambda {el:.2e}", flush=True)
    assert ex < 1e-4 and el < 1e-4
