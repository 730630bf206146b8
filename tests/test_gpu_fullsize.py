"""GPU parity at full workload sizes.

The GPU denoises the whole volume; the CPU oracle then recomputes, for a
small window of voxels, every selected patch covering the window and the
reference's recombination (weighted: theta = 1/(2+r), out / w) on those
voxels only.  The window output and weights must match, and the retained rank
of every selected patch must equal the oracle's (computed by a CPU process
pool, oracle/pool.py) except at ties: where the oracle's own decision
margin is within float32 rounding of the patch's largest eigenvalue.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _window_reference(data, mask, ps, po, patch_fn, sel, lo, hi):
    """Oracle recombination restricted to voxels lo..hi (spatial, inclusive-exclusive)."""
    from oracle import cpu_ref as O

    src = O.StridedPatches(data, ps, po)
    grid = src.grid
    step = [p - o for p, o in zip(ps, po)]
    ws = tuple(h - l for l, h in zip(lo, hi))
    out = np.zeros(ws + (data.shape[-1],), np.complex128 if np.iscomplexobj(data) else np.float64)
    w = np.zeros(ws + (data.shape[-1],), np.float64)
    M = int(np.prod(ps[:-1]))
    ranks = {}
    for L in sel:
        g = np.unravel_index(L, grid)
        o = [gi * st for gi, st in zip(g[:3], step[:3])]
        if any(o[d] + ps[d] <= lo[d] or o[d] >= hi[d] for d in range(3)):
            continue
        P, r, _ = patch_fn(src.get(L).reshape(M, -1))
        P = P.reshape(ps)
        ranks[int(L)] = int(r)
        th = 1.0 / (2 + r)
        a = [max(lo[d], o[d]) for d in range(3)]
        b = [min(hi[d], o[d] + ps[d]) for d in range(3)]
        sl_out = tuple(slice(a[d] - lo[d], b[d] - lo[d]) for d in range(3))
        sl_p = tuple(slice(a[d] - o[d], b[d] - o[d]) for d in range(3))
        out[sl_out] += th * P[sl_p]
        w[sl_out] += th
    with np.errstate(invalid="ignore"):
        res = out / w
    return res, w, ranks


def _run_gpu(denoiser, data, mask):
    from patch_denoise import _native as N

    prob, p_s, p_o, mask_arr = denoiser._prepare(data, mask, 50)
    var_map = denoiser._configure(prob, data, p_s)
    run = N.DeviceRun(prob, data, mask_arr, var_map, want_patch_info=True).run()
    out, w = run.result(data.shape)
    return out, w, run.patch_info()


def test_mppca_80x80x48x300_window_and_ranks():
    """Headline workload: MP-PCA, p=5, stride 1, ellipsoid mask, weighted."""
    from oracle import cpu_ref as O
    from patch_denoise import MPPCADenoiser, workloads

    data, mask = workloads.config2()
    ps, po = (5, 5, 5, 300), (4, 4, 4, 300)
    den = MPPCADenoiser(5, 4, threshold_scale=1.0)
    print("[fullsize] mppca 80x80x48x300 gpu run", flush=True)
    out, w, info = _run_gpu(den, data, mask)
    sel = info["index"].astype(np.int64)
    assert len(sel) == 107680
    # every selected patch's rank against the oracle (CPU process pool); a
    # mismatch is allowed only at a tie: the oracle's own decision margin
    # within float32 rounding of the largest eigenvalue
    from oracle import pool as OP

    with OP.SharedVolume(data) as sv:
        r_or, margin = OP.oracle_ranks(sv, ps, po, "mppca", {"ts": 1.0}, sel)
    bad = np.nonzero(r_or != info["rank"])[0]
    print(f"[fullsize] rank mismatches {len(bad)} / {len(sel)} (all patches), margins "
          f"{[f'{margin[b]:.1e}' for b in bad[:20]]}", flush=True)
    assert np.all(margin[bad] < 1e-5), "rank mismatch away from the cutoff"
    assert len(bad) <= 20
    # a 4x4x4-voxel window at the mask boundary and one inside
    for lo in ((38, 38, 22), (6, 38, 22)):
        hi = tuple(v + 4 for v in lo)
        ref, ref_w, _ = _window_reference(data, mask, ps, po,
                                          lambda X: O.mppca_patch(X, 1.0), sel, lo, hi)
        win = out[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        m = mask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        ref = np.where(m[..., None], ref, 0.0)
        e = rel_l2(np.nan_to_num(win), np.nan_to_num(ref))
        ew = rel_l2(w[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]], ref_w)
        print(f"[fullsize] window {lo}: rel-L2 out {e:.2e} weights {ew:.2e}", flush=True)
        assert e < 1e-5 and ew < 1e-6


def test_optimal_fro_64x64x32x120_window():
    """Optimal (Frobenius) shrinkage on the 64x64x32x120 rank-6 workload, p=5 o=4."""
    from oracle import cpu_ref as O
    from patch_denoise import OptimalSVDDenoiser, workloads

    data, clean = workloads.config5()
    ps, po = (5, 5, 5, 120), (4, 4, 4, 120)
    den = OptimalSVDDenoiser(5, 4, loss="fro")
    mpm = O.mp_median(120 / 125)
    den.input_denoising_kwargs["mp_median"] = mpm
    den._var_map = None
    print("[fullsize] optimal-fro 64x64x32x120 gpu run", flush=True)
    out, w, info = _run_gpu(den, data, None)
    sel = info["index"].astype(np.int64)
    from oracle import pool as OP

    with OP.SharedVolume(data) as sv:
        r_or, margin = OP.oracle_ranks(sv, ps, po, "optimal",
                                       {"loss": "fro", "mpm": mpm, "var": None}, sel)
    bad = np.nonzero(r_or != info["rank"])[0]
    print(f"[fullsize] rank mismatches {len(bad)} / {len(sel)} (all patches), margins "
          f"{[f'{margin[b]:.1e}' for b in bad[:20]]}", flush=True)
    assert np.all(margin[bad] < 1e-5), "rank mismatch away from the cutoff"
    lo = (30, 30, 14)
    hi = tuple(v + 3 for v in lo)
    ref, ref_w, _ = _window_reference(
        data, None, ps, po, lambda X: O.optimal_patch(X, O.shrink_fro, mpm), sel, lo, hi)
    win = out[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    e = rel_l2(win, ref)
    print(f"[fullsize] window rel-L2 out {e:.2e}", flush=True)
    assert e < 1e-5
    # No "denoising helps" check here: at beta = T/M = 0.96 the reference's
    # ungated Frobenius shrinker (lowrank.py:272-280) amplifies singular values
    # below the lower bulk edge 1 - sqrt(beta), and the output reproduces that.


def test_nordic_complex_window():
    """NORDIC on complex64 (reduced 32x32x20x100 volume of the complex workload
    recipe, p=5: Gram size min(125, 100) = 100)."""
    from oracle import cpu_ref as O
    from patch_denoise import workloads
    import patch_denoise as pd

    data = workloads.config3(shape=(32, 32, 20), T=100)
    ps, po = (5, 5, 5, 100), (4, 4, 4, 100)
    np.random.seed(0)
    print("[fullsize] nordic complex 32x32x20x100 gpu run", flush=True)
    out, w, _, _ = pd.nordic(data, 5, 4, noise_std=25.0, n_iter_threshold=3, progbar=False)
    np.random.seed(0)
    tau = O.nordic_threshold(ps, 100, 25.0, 3)
    lo = (14, 14, 8)
    hi = tuple(v + 3 for v in lo)
    src = O.StridedPatches(data, ps, po)
    sel = np.arange(src.n_patches)
    ref, ref_w, _ = _window_reference(data, None, ps, po, lambda X: O.raw_patch(X, tau), sel,
                                      lo, hi)
    win = out[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    e = rel_l2(win, ref)
    print(f"[fullsize] window rel-L2 out {e:.2e}", flush=True)
    assert e < 1e-5


def test_split_bf16_gram_mppca():
    """Split-BF16 three-product Gram (precision="bf16x3"): the 80x80x48x300
    MP-PCA output within 1e-3 relative L2 of the FP32 path (and of the oracle
    on a window); ranks equal except near-ties."""
    from oracle import cpu_ref as O
    from patch_denoise import MPPCADenoiser, workloads

    data, mask = workloads.config2()
    ps, po = (5, 5, 5, 300), (4, 4, 4, 300)
    print("[fullsize] mppca 80x80x48x300 fp32 vs bf16x3", flush=True)
    out32, w32, info32 = _run_gpu(MPPCADenoiser(5, 4, threshold_scale=1.0, precision="fp32"), data, mask)
    outbf, wbf, infobf = _run_gpu(MPPCADenoiser(5, 4, threshold_scale=1.0, precision="bf16x3"), data,
                                  mask)
    m = mask[..., None] & np.isfinite(out32)
    e = rel_l2(outbf[np.broadcast_to(m, out32.shape)], out32[np.broadcast_to(m, out32.shape)])
    nd = int((info32["rank"] != infobf["rank"]).sum())
    print(f"[fullsize] bf16x3 vs fp32: rel-L2 {e:.2e}, rank differences {nd} / {len(info32['rank'])}",
          flush=True)
    assert e < 1e-3
    assert nd <= len(info32["rank"]) // 100
    sel = infobf["index"].astype(np.int64)
    lo = (38, 38, 22)
    hi = tuple(v + 4 for v in lo)
    ref, _, _ = _window_reference(data, mask, ps, po, lambda X: O.mppca_patch(X, 1.0), sel, lo, hi)
    win = outbf[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    ref = np.where(mask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]][..., None], ref, 0.0)
    ew = rel_l2(np.nan_to_num(win), np.nan_to_num(ref))
    print(f"[fullsize] bf16x3 window vs oracle {ew:.2e}", flush=True)
    assert ew < 1e-3
