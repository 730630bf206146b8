"""Benchmark: MP-PCA voxel-frames/s on the 80x80x48x300 workload, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-stride S] [--no-cpu]

A step is one full ``pdamd_denoise`` call over the volume (selection, Gram +
tridiagonalisation, eigen/threshold, synthesis + recombination,
normalisation) with the input already resident in HBM; the timed region is
bracketed by ``torch.cuda.synchronize()``.  Per-stage kernel times are taken
live from HIP events recorded on the work stream during the timed steps.

Prints one JSON line with ``roofline`` (dominant kernel) and ``cpu_baseline``
(the CPU oracle, single-threaded, on a strided sample of the same patches,
extrapolated to the full volume).
"""

import argparse
import ctypes
import json
import os
import sys
import time

os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
os.environ.setdefault("OMP_NUM_THREADS", "1")

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "patch-denoising_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
WORKLOAD = "mppca_80x80x48x300_p5_o4_ellipsoid_weighted"


def algorithmic_work(n, N, M, T, ranks):
    """Algorithmic flops of each per-patch kernel summed over the selected patches.

    gram:    n(n+1)N real flops per patch (SYRK count of the min-side Gram).
    tridiag: 4/3 n^3 per patch (Householder tridiagonalisation).
    synth:   4 M T r per patch (two rank-r products).
    """
    P = len(ranks)
    return dict(gram=P * n * (n + 1) * N, tridiag=P * 4.0 / 3.0 * n**3,
                synth=float(4.0 * M * T * np.sum(ranks)))


def cpu_baseline(data, mask, ps, po, sel, stride):
    """Time the CPU oracle (single thread) on every ``stride``-th selected patch."""
    from oracle import cpu_ref as O

    src = O.StridedPatches(data, ps, po)
    out = np.zeros(data.shape, np.float32)
    w = np.zeros(data.shape, np.float32)
    dst, wts = O.StridedPatches(out, ps, po), O.StridedPatches(w, ps, po)
    sample = sel[::stride]
    M = int(np.prod(ps[:-1]))
    from threadpoolctl import threadpool_info, threadpool_limits

    with threadpool_limits(limits=1):
        blas = [(d.get("internal_api"), d.get("num_threads")) for d in threadpool_info()]
        t0 = time.perf_counter()
        for i in sample:
            X = src.get(i).reshape(M, -1)
            P, r, _ = O.mppca_patch(X, 1.0)
            th = 1 / (2 + r)
            o = dst.get(i)
            o += P.reshape(ps) * th
            ww = wts.get(i)
            ww += th
        dt = time.perf_counter() - t0
    per_patch = dt / len(sample)
    full = per_patch * len(sel)
    return dict(sample_patches=int(len(sample)), seconds=dt, per_patch_s=per_patch,
                full_estimate_s=full, blas=blas)


def _mp_rank_torch(a, M, T, ts=1.0):
    """MP-PCA rank from the ascending top-10 eigenvalues ``a`` [B, 10] of the
    M x M covariance, without host synchronisation: the rule of
    lowrank.py:42-63 unrolled over its at most 10 loop steps."""
    import torch

    B, K = a.shape
    k = torch.zeros(B, dtype=torch.long, device=a.device)
    active = torch.ones(B, dtype=torch.bool, device=a.device)
    cs = torch.cumsum(a, 1)
    for step in range(K):
        if step == 0:
            var = cs[:, -1] / K * (4.0 * ((K + 1) / M) ** 0.5)
        else:
            var = cs[:, K - 1 - step] / (K - step) * (4.0 * ((T - step + 1) / M) ** 0.5)
        go = active & (var < a[:, K - 1 - step] - a[:, 0])
        k = k + go.long()
        active = go
    left = (K - k).clamp(min=1)
    noise = torch.gather(cs, 1, (left - 1)[:, None])[:, 0] / left
    noise = torch.where(k < K, noise, torch.full_like(noise, float("nan")))
    return (a > (noise * ts * ts)[:, None]).sum(1)


def _patch_voxels(sel, shape, p=5):
    """Flat voxel indices [P, p^3] of stride-1 patches with linear grid indices ``sel``."""
    import torch

    X, Y, Z = shape[:3]
    gy, gz = Y - p + 1, Z - p + 1
    i, j, k = sel // (gy * gz), (sel // gz) % gy, sel % gz
    o = torch.arange(p, device=sel.device)
    off = ((o[:, None, None] * Y + o[None, :, None]) * Z + o[None, None, :]).reshape(-1)
    return ((i * Y + j) * Z + k)[:, None] + off[None, :]


def _lowrank_torch(xf, vidx, M, T, out, w):
    """Gather, centre, Gram, batched torch.linalg.eigh (FP32), MP-PCA rule,
    projection onto the kept eigenvectors, weighted scatter-add (index_add_)."""
    import torch

    Xp = xf[vidx]                                    # [B, M, T]
    m = Xp.mean(1, keepdim=True)
    Xc = Xp - m
    G = torch.matmul(Xc, Xc.transpose(1, 2).contiguous()) / (M - 1)
    ev, V = torch.linalg.eigh(G)
    r = _mp_rank_torch(ev[:, -10:], M, T)
    keep = (torch.arange(M, device=xf.device)[None, :] >= (M - r)[:, None]).to(V.dtype)
    Vk = V * keep[:, None, :]
    Pd = torch.matmul(Vk, torch.matmul(Vk.transpose(1, 2).contiguous(), Xc)) + m
    th = (1.0 / (2 + r)).to(xf.dtype)
    out.index_add_(0, vidx.reshape(-1), (th[:, None, None] * Pd).reshape(-1, T))
    w.index_add_(0, vidx.reshape(-1), th[:, None].expand(-1, M).reshape(-1, 1).expand(-1, T))
    return r


def plain_gpu_baseline(data, sel, stride, M=125):
    """The plain one-patch-per-launch FP32 GPU path: per selected patch, a
    chain of PyTorch launches (gather, centre, Gram matmul, torch.linalg.eigh,
    MP-PCA rule, projection, scatter-add) with no batching and no host
    synchronisation.  Timed on every ``stride``-th selected patch and
    extrapolated to all of them."""
    import torch

    x = torch.from_numpy(data).cuda()
    T = data.shape[-1]
    xf = x.reshape(-1, T)
    out = torch.zeros_like(xf)
    w = torch.zeros_like(xf)
    sample = torch.from_numpy(np.asarray(sel[::stride], np.int64)).cuda()
    vids = _patch_voxels(sample, data.shape)
    for b in range(min(20, len(sample))):
        _lowrank_torch(xf, vids[b:b + 1], M, T, out, w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in range(len(sample)):
        _lowrank_torch(xf, vids[b:b + 1], M, T, out, w)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    per = dt / len(sample)
    return dict(per_patch_ms=per * 1e3, sample_patches=int(len(sample)), seconds=dt,
                full_estimate_s=per * len(sel))


def batched_gpu_baseline(data, sel, ranks_ref, batch=4096, budget_s=20.0, M=125):
    """The reference's own GPU-path shape (gpu/denoiser.py:64-96: a batch of
    patches through one batched decomposition, then one fold) in plain
    PyTorch FP32: per batch of ``batch`` selected patches, a gather, a
    batched Gram, a batched ``torch.linalg.eigh``, the MP-PCA rule, the
    projection and a weighted ``index_add_``.  Batches run in order until
    ``budget_s`` has passed; the time is extrapolated to all patches.  The
    ranks of the batches run are compared with ``ranks_ref``."""
    import torch

    x = torch.from_numpy(data).cuda()
    T = data.shape[-1]
    xf = x.reshape(-1, T)
    out = torch.zeros_like(xf)
    w = torch.zeros_like(xf)
    sel_t = torch.from_numpy(np.asarray(sel, np.int64)).cuda()
    _lowrank_torch(xf, _patch_voxels(sel_t[:batch], data.shape), M, T, out, w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, agree = 0, 0
    rs = []
    while done < len(sel):
        b = sel_t[done:done + batch]
        rs.append(_lowrank_torch(xf, _patch_voxels(b, data.shape), M, T, out, w))
        done += len(b)
        torch.cuda.synchronize()
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    r = torch.cat(rs).cpu().numpy()
    agree = float(np.mean(r == np.asarray(ranks_ref[:done])))
    return dict(per_patch_ms=dt / done * 1e3, sample_patches=int(done), seconds=dt,
                full_estimate_s=dt / done * len(sel), batch=batch,
                rank_agreement_with_hip=agree)


def bench_differentiable(steps):
    """Forward and forward+backward of the differentiable optimal-shrinkage
    path on the rank-6 64x64x32x120 workload (patch 5, overlap 4, Frobenius
    loss, per-patch lambda = 1, loss = MSE to the clean signal)."""
    import torch

    from patch_denoise import _native as N
    from patch_denoise import workloads
    from patch_denoise.differentiable import _Spec, optimal_shrinkage

    noisy, clean = workloads.config5()
    x = torch.from_numpy(noisy).cuda().requires_grad_(True)
    tgt = torch.from_numpy(clean).cuda()
    spec = _Spec(noisy.shape, np.float32, 5, 4)
    lam = torch.ones(spec.grid4[:3], dtype=torch.float32, device="cuda", requires_grad=True)

    def fwd():
        return optimal_shrinkage(x, lam, 5, 4, spec=spec)

    def step():
        x.grad = None
        lam.grad = None
        loss = ((fwd() - tgt) ** 2).mean()
        loss.backward()

    with torch.no_grad():
        fwd()
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        for _ in range(steps):
            fwd()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / steps
    # per-stage breakdown from separate steps (the HIP-event timers add a
    # synchronisation per call, so they stay out of the timed loop)
    ms = (ctypes.c_float * len(N.STAGES))()
    cnt = (ctypes.c_int64 * len(N.STAGES))()
    spec.timers = (ms, cnt)
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    spec.timers = None
    ok = bool(torch.isfinite(x.grad).all() and torch.isfinite(lam.grad).all())
    vf = float(np.prod(noisy.shape))
    return dict(workload="optimal_fro_64x64x32x120_p5_o4_lambda_mse", dtype_in="f32",
                eigen_dtype="f64", patches=spec.selected, steps=steps,
                fwd_ms=t_f * 1e3, fwd_bwd_ms=t_fb * 1e3, ratio=t_fb / t_f,
                fwd_voxel_frames_s=vf / t_f, fwd_bwd_voxel_frames_s=vf / t_fb,
                grads_finite=ok,
                stage_ms_per_step={name: ms[i] / steps for i, name in enumerate(N.STAGES)
                                   if cnt[i]})


def bench_batched(steps, autotune=True):
    """The eight-volume mixed workload (seed 4, all eight volumes, Gram sizes
    27 to 200): one pdamd_denoise call per volume vs
    one pdamd_denoise_batch call (eager, and replayed from its captured
    graph), inputs resident in HBM."""
    import torch

    from patch_denoise import _native as N
    from patch_denoise import workloads
    from patch_denoise.batch import BatchPlan

    t0 = time.perf_counter()
    vols = workloads.config4()
    calls, skipped = workloads.config4_calls(vols)
    gen_s = time.perf_counter() - t0
    # at most two sets of workspaces alive at a time (the single calls and the
    # eager batch, then the eager batch and the graph)
    eager = BatchPlan(calls, use_graph=False)
    singles = [N.DeviceRun(j.prob, j.data, j.mask_arr, j.var_map) for j in eager.jobs]
    vf = float(sum(np.prod(j.data.shape) for j in eager.jobs))

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    t_seq = timed(lambda: [r.run() for r in singles])
    del singles
    t_eager = timed(eager.run)
    graph = BatchPlan(calls, use_graph=True)
    t_graph = timed(graph.run)
    del graph
    # error of every volume against the same call on its float64 cast (the
    # FP64 path end to end)
    from patch_denoise import denoise as F

    outs32 = eager.results()
    errs = []
    for (name, args, kw), o32 in zip(calls, outs32):
        o64 = getattr(F, name)(args[0].astype(np.float64), *args[1:], **kw)
        errs.append(_rel_l2(np.asarray(o32[0]), np.asarray(o64[0])))
        del o64
    tuned_res = {}
    if autotune:
        t0 = time.perf_counter()
        tuned = BatchPlan(calls, use_graph=True, autotune=True)
        tune_s = time.perf_counter() - t0
        t_tuned = timed(tuned.run)
        tuned_res = dict(autotuned_graph_ms=t_tuned * 1e3, autotune_s=tune_s,
                         autotuned_launch_per_cu=[list(t) for t in tuned.tuned],
                         speedup_autotuned_vs_sequential=t_seq / t_tuned)
    return dict(workload="batched_8vol_seed4", volumes=[
        dict(method=c[0], shape=list(c[1][0].shape), patch=c[1][1]) for c in calls],
        skipped=[dict(index=i, reason=r) for i, r in skipped], voxel_frames=vf,
        patches=eager.batch.selected, steps=steps, datagen_s=gen_s,
        rel_l2_vs_fp64=errs,
        sequential_ms=t_seq * 1e3, batch_ms=t_eager * 1e3, batch_graph_ms=t_graph * 1e3,
        sequential_voxel_frames_s=vf / t_seq, batch_graph_voxel_frames_s=vf / t_graph,
        speedup_graph_vs_sequential=t_seq / t_graph, **tuned_res)


def bench_large(steps):
    """Gram sizes above 128 (global-memory + packed-LDS Householder): the
    49x49x52x200 float32 volume of the batched workload (p = 7, MP-PCA, Gram
    200) and the complex64 64x64x40x200 NORDIC workload (p = 11, Gram 200,
    noise_std 25) at its default precision (FP64 eigen stage), with an FP32
    eigen stage and with the split-BF16 Gram, one pdamd_denoise call each,
    inputs resident in HBM (NORDIC's host threshold draws run once, untimed)."""
    import ctypes

    import torch

    from patch_denoise import _native as N
    from patch_denoise import workloads
    from patch_denoise.batch import BatchPlan

    vol4 = workloads.config4_calls(workloads.config4())[0][4]
    c3 = workloads.config3()
    cases = [("mppca_49x49x52x200_p7_o6_ellipsoid_weighted", vol4),
             ("nordic_c64_64x64x40x200_p11_o10_default",
              ("nordic", (c3, 11, 10), {"noise_std": 25.0})),
             ("nordic_c64_64x64x40x200_p11_o10_fp32",
              ("nordic", (c3, 11, 10), {"noise_std": 25.0, "precision": "fp32"})),
             ("nordic_c64_64x64x40x200_p11_o10_bf16x3",
              ("nordic", (c3, 11, 10), {"noise_std": 25.0, "precision": "bf16x3"}))]
    res = {}
    outs = {}
    for name, call in cases:
        job = BatchPlan([call], use_graph=False).jobs[0]
        run = N.DeviceRun(job.prob, job.data, job.mask_arr, job.var_map)
        run.run()
        torch.cuda.synchronize()
        ms = (ctypes.c_float * len(N.STAGES))()
        cnt = (ctypes.c_int64 * len(N.STAGES))()
        job.prob.stage_ms = ctypes.cast(ms, ctypes.c_void_p).value
        job.prob.stage_launches = ctypes.cast(cnt, ctypes.c_void_p).value
        t0 = time.perf_counter()
        for _ in range(steps):
            run.run()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        job.prob.stage_ms = None
        job.prob.stage_launches = None
        vf = float(np.prod(job.data.shape))
        n = min(int(np.prod(job.prob.patch[:3])), int(job.prob.patch[3]))
        res[name] = dict(gram_size=n, patch=list(job.prob.patch), ms_per_call=t * 1e3,
                         voxel_frames_s=vf / t, steps=steps,
                         stage_ms_per_call={nm: ms[i] / steps for i, nm in enumerate(N.STAGES)
                                            if cnt[i]})
        if name.startswith("nordic"):
            outs[name] = run.out.clone()
        del run
    base = outs["nordic_c64_64x64x40x200_p11_o10_default"]
    for nm in ("fp32", "bf16x3"):
        res[f"nordic_c64_64x64x40x200_p11_o10_{nm}"]["rel_l2_vs_default"] = _rel_l2(
            outs[f"nordic_c64_64x64x40x200_p11_o10_{nm}"], base)
    del outs
    return res


def bench_noise_map(steps):
    """Sliding-window noise map (estimate_noise) on an 80x80x48x300 float32
    volume, block 3, input resident in HBM.  HBM-bound: the algorithmic bytes
    are one read of the volume plus the per-voxel moments and the map."""
    import torch

    from patch_denoise import _native as N

    g = torch.Generator(device="cuda").manual_seed(8)
    x = 800.0 + 30.0 * torch.randn((80, 80, 48, 300), generator=g, device="cuda",
                                   dtype=torch.float32)
    N.estimate_noise(x, 3)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        N.estimate_noise(x, 3)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nvox = 80 * 80 * 48
    algo = x.numel() * 4 + nvox * (3 * 8 * 2 + 8)
    return dict(workload="noise_map_80x80x48x300_block3", ms_per_call=ms,
                algorithmic_bytes=algo, achieved_GBs=algo / (ms / 1e3) / 1e9,
                peak_GBs=HBM_PEAK_GBS, frac=algo / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                note="time includes the two kernels and the per-call buffer setup")


def _rel_l2(a, b):
    """Relative L2 distance over the elements finite in both (complex-safe)."""
    import torch

    a, b = torch.as_tensor(a), torch.as_tensor(b)
    fin = torch.isfinite(a) & torch.isfinite(b)
    d = (a.to(b.dtype) - b)[fin].abs().double()
    return float(d.pow(2).sum().sqrt() / b[fin].abs().double().pow(2).sum().sqrt().clamp_min(1e-300))


def fp64_headline(dev_x, s4, p4, st4, mask, out32, ranks32):
    """The headline call on the float64 cast of the input (float64 Gram,
    eigen stage and accumulation) against the float32 default path."""
    import torch

    from patch_denoise import _native as N

    prob = N.make_problem(s4, p4, st4, np.float64, N.MPPCA, "weighted", "auto")
    prob.threshold_scale2 = 1.0
    prob.scale_f64 = 0
    prob.mask_kind = N.MASK_SPATIAL
    prob.mask_threshold = 50.0
    run = N.DeviceRun(prob, dev_x.double(), mask, None, want_patch_info=True).run()
    torch.cuda.synchronize()
    r64 = run.patch_info()["rank"]
    res = dict(rel_l2_vs_fp64=_rel_l2(out32, run.out),
               rank_mismatches_vs_fp64=int(np.sum(np.asarray(r64) != np.asarray(ranks32))),
               patches=int(len(r64)), out64=run.out)
    return res


def bench_public_calls(steps):
    """NumPy-in / NumPy-out calls of the public API (host arrays in, host
    arrays out: the host-to-device copy, the call and the copy back are all
    timed), and each one's rel-L2 against the same call on the float64 (or
    complex128) cast of its input:

    * MP-PCA 80x80x48x300, p=5, o=4, ellipsoid mask, weighted (``mp_pca``);
    * NORDIC complex64 64x64x40x200, p=11, o=10, noise_std 25, default
      precision (``nordic``; ``np.random.seed(0)`` before every call);
    * optimal shrinkage (Frobenius) float32 64x64x32x120, p=5, o=4
      (``optimal_thresholding``)."""
    import torch

    from patch_denoise import mp_pca, nordic, optimal_thresholding, workloads

    d2, m2 = workloads.config2()
    d3 = workloads.config3()
    d5, _ = workloads.config5()

    def c2(x):
        return mp_pca(x, 5, 4, mask=m2, threshold_scale=1.0)

    def c3(x):
        np.random.seed(0)
        return nordic(x, 11, 10, noise_std=25.0)

    def c5(x):
        return optimal_thresholding(x, 5, 4, loss="fro")

    res = {}
    for name, fn, x, x64 in [
            ("mppca_80x80x48x300_p5_o4_ellipsoid_weighted", c2, d2, d2.astype(np.float64)),
            ("nordic_c64_64x64x40x200_p11_o10_default", c3, d3, d3.astype(np.complex128)),
            ("optimal_fro_64x64x32x120_p5_o4", c5, d5, d5.astype(np.float64))]:
        out = fn(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = fn(x)
        t = (time.perf_counter() - t0) / steps
        ref = fn(x64)
        res[name] = dict(ms_per_call=t * 1e3, voxel_frames_s=float(x.size) / t, steps=steps,
                         dtype=str(x.dtype), rel_l2_vs_fp64=_rel_l2(
                             torch.from_numpy(np.asarray(out[0])),
                             torch.from_numpy(np.asarray(ref[0]))))
        del out, ref
    return res


def cpu_pool_baseline(data, ps, po, sel, stride):
    """The oracle's MP-PCA on a pool of one process per host core (at most
    16), one BLAS thread each, over every ``stride``-th selected patch."""
    from oracle import pool as OP

    sample = np.asarray(sel[::stride], np.int64)
    with OP.SharedVolume(data) as vol:
        r = OP.time_mppca_pool(vol, ps, po, 1.0, sample)
    per = r["wall_s"] / len(sample)
    return dict(sample_patches=int(len(sample)), seconds=r["wall_s"], busy_s=r["busy_s"],
                workers=r["workers"], per_patch_s=per, full_estimate_s=per * len(sel))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-stride", type=int, default=32,
                    help="single-core CPU baseline: every S-th selected patch")
    ap.add_argument("--pool-stride", type=int, default=8,
                    help="CPU pool baseline: every S-th selected patch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain one-patch-per-launch GPU baseline")
    ap.add_argument("--plain-stride", type=int, default=64)
    ap.add_argument("--precision", default="auto")
    ap.add_argument("--batch-steps", type=int, default=2,
                    help="steps of the batched multi-volume measurement; 0 skips it")
    ap.add_argument("--diff-steps", type=int, default=3,
                    help="steps of the differentiable (fwd / fwd+bwd) measurement; 0 skips it")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the batched leg's autotuned measurement (for kernel traces)")
    ap.add_argument("--large-steps", type=int, default=2,
                    help="steps of the Gram-size-200 measurements; 0 skips them")
    ap.add_argument("--only", choices=["all", "headline", "differentiable", "batched", "noise",
                                       "large"],
                    default="all", help="run only one part (for profiling)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: patches would "
              "shard across GPUs by volume region, which is not built.", file=sys.stderr)
        sys.exit(2)

    import torch

    from patch_denoise import _native as N
    from patch_denoise import workloads
    from patch_denoise.space_time.base import _geometry

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.only == "differentiable":
        print(json.dumps({"differentiable": bench_differentiable(max(args.diff_steps, 1))}))
        return
    if args.only == "batched":
        print(json.dumps({"batched": bench_batched(max(args.batch_steps, 1),
                                                   autotune=not args.no_autotune)}))
        return
    if args.only == "noise":
        print(json.dumps({"noise_map": bench_noise_map(5)}))
        return
    if args.only == "large":
        print(json.dumps({"large_gram": bench_large(max(args.large_steps, 1))}))
        return
    t_gen = time.perf_counter()
    data, mask = workloads.config2()
    gen_s = time.perf_counter() - t_gen
    shape = data.shape
    ps, po = (5, 5, 5, shape[-1]), (4, 4, 4, shape[-1])
    s4, p4, st4 = _geometry(shape, ps, po)
    prob = N.make_problem(s4, p4, st4, np.float32, N.MPPCA, "weighted", args.precision)
    prob.threshold_scale2 = 1.0
    prob.scale_f64 = 0
    prob.mask_kind = N.MASK_SPATIAL
    prob.mask_threshold = 50.0
    dev_x = torch.from_numpy(data).cuda()
    run = N.DeviceRun(prob, dev_x, mask, None, want_patch_info=True)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        run.run()
    torch.cuda.synchronize()

    stage_ms = (ctypes.c_float * len(N.STAGES))()
    stage_n = (ctypes.c_int64 * len(N.STAGES))()
    prob.stage_ms = ctypes.cast(stage_ms, ctypes.c_void_p).value
    prob.stage_launches = ctypes.cast(stage_n, ctypes.c_void_p).value
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.run()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prob.stage_ms = None
    prob.stage_launches = None

    if args.calib_copy:
        # plain copy kernel of known size: 256 MiB read + 256 MiB written
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ms_per_step = elapsed * 1e3 / args.steps
    vf = float(np.prod(shape))
    value = vf / (ms_per_step / 1e3)
    info = run.patch_info()
    ranks = np.asarray(info["rank"])
    sel = np.asarray(info["index"])
    M, T = 125, shape[-1]
    n, Nn = min(M, T), max(M, T)
    P = len(ranks)
    K = 10
    nvox = int(np.prod(shape[:3]))
    npatch = int(np.prod([shape[d] - 4 for d in range(3)]))
    stages = {name: dict(ms_per_step=stage_ms[i] / args.steps,
                         launches_per_step=stage_n[i] / args.steps)
              for i, name in enumerate(N.STAGES) if stage_n[i]}

    # Roofline of the dominant kernel, the Gram (gather + centre + SYRK on
    # MFMA): algorithmic flops n(n+1)N per selected patch over its average
    # launch duration (HIP events on the work stream).
    f_gram = float(P) * n * (n + 1) * Nn
    f_syn = 4.0 * M * T * float(np.sum(ranks))
    gram_peak = FP32_PEAK_TFLOPS if args.precision != "bf16x3" else BF16_PEAK_TFLOPS / 3
    g = stages["gram"]
    g_launch_ms = g["ms_per_step"] / g["launches_per_step"]
    g_ach = f_gram / g["launches_per_step"] / (g_launch_ms / 1e3) / 1e12
    roof = dict(bound="mfma", achieved=g_ach, peak=gram_peak, unit="TFLOP/s",
                frac=g_ach / gram_peak, traffic=None, kernel="gram_kernel",
                avg_launch_ms=g_launch_ms, launches_per_step=g["launches_per_step"],
                flops_per_launch=f_gram / g["launches_per_step"],
                work=f"n(n+1)N flops per selected patch (n={n}, N={Nn}), {P} patches per step")

    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # summary of this workload (each scaled by a copy kernel of known size).
    tr = {}
    tpath = os.path.join(REPO, "profiles", f"hbm_traffic_{WORKLOAD.split('_p5')[0]}.json")
    if os.path.exists(tpath):
        tr = json.load(open(tpath))
        k = tr.get("kernels", {}).get(roof["kernel"])
        if k is not None:
            roof["traffic"] = k["bytes_per_dispatch"]
            roof["traffic_source"] = tr["source"]

    # The other kernels of the step against their own roofs.
    covered = int((run.w[..., 0] > 0).sum())
    kernels = {}
    if "recombine" in stages:
        fac_b = covered * T * 4 + P * (T * 4 + n * K * 4 + K * 8 + (K + 1) * T * 4)
        rec_b = P * ((K + 1) * T * 4 + n * K * 4) + npatch * 4 + nvox * 5 + 2 * nvox * T * 4
        for name, stage, kern, byts, flops, what in [
                ("gather_factors", "synth", "gather_factor_kernel", fac_b, 2.0 * M * T * K * P,
                 "reads each selected patch from the volume (the Casorati gather), its "
                 "per-frame mean and kept eigenvectors; writes theta*[mean; wc*U^T Xc] "
                 "(11 x T) per patch"),
                ("recombine", "recombine", "gather_recombine_kernel", rec_b,
                 2.0 * M * T * (K + 1) * P,
                 "sums every output element over its covering patches' factors in "
                 "patch order; reads the factors once, writes output and weights")]:
            ms = stages[stage]["ms_per_step"]
            gbs = byts / (ms / 1e3) / 1e9
            tfl = flops / (ms / 1e3) / 1e12
            ent = dict(kernel=kern, ms_per_step=ms, algorithmic_bytes=byts, achieved_GBs=gbs,
                       hbm_peak_GBs=HBM_PEAK_GBS, hbm_frac=gbs / HBM_PEAK_GBS, flops=flops,
                       achieved_TFLOPs=tfl, valu_peak_TFLOPs=FP32_PEAK_TFLOPS,
                       valu_frac=tfl / FP32_PEAK_TFLOPS, what=what)
            kt = tr.get("kernels", {}).get(kern)
            if kt is not None:
                ent["traffic"] = kt["bytes_per_dispatch"] * kt.get("dispatches", 1)
            kernels[name] = ent
    eig_ms = sum(stages.get(k, {}).get("ms_per_step", 0.0) for k in ("tridiag", "trieig"))
    kernels["eigensolve"] = dict(
        ms_per_step=eig_ms, share_of_step=eig_ms / ms_per_step,
        note="Householder tridiagonalisation + bisection / twisted-factorisation "
             "eigenvectors + MP-PCA rule: VALU/LDS latency work, kept out of the roofline")
    kernels["end_to_end"] = dict(
        flops=f_gram + f_syn, achieved_TFLOPs=(f_gram + f_syn) / (ms_per_step / 1e3) / 1e12,
        peak=gram_peak, frac=(f_gram + f_syn) / (ms_per_step / 1e3) / 1e12 / gram_peak,
        work="sum over selected patches of n(n+1)N + 4 M T r (r = kept rank)")

    cpu = None
    if not args.no_cpu:
        pool = cpu_pool_baseline(data, ps, po, sel, args.pool_stride)
        cb = cpu_baseline(data, mask, ps, po, sel, args.cpu_stride)
        cpu = dict(value=vf / pool["full_estimate_s"], unit="voxel-frames/s",
                   cores=pool["workers"], kind="port",
                   sample=(f"oracle MP-PCA (weighted accumulation) on every {args.pool_stride}th "
                           f"selected patch ({pool['sample_patches']} of {len(sel)}), one process"
                           f" per core with one BLAS thread each, {pool['seconds']:.1f} s wall, "
                           "extrapolated to all selected patches"),
                   per_patch_ms=pool["per_patch_s"] * 1e3, host_cpus=os.cpu_count(),
                   single_core=dict(
                       value=vf / cb["full_estimate_s"], cores=1,
                       sample=(f"every {args.cpu_stride}th selected patch "
                               f"({cb['sample_patches']} patches, {cb['seconds']:.1f} s)"),
                       per_patch_ms=cb["per_patch_s"] * 1e3, blas_threads=cb["blas"]))
        cpu["pool_speedup_vs_single_core"] = cpu["value"] / cpu["single_core"]["value"]

    prec = None
    if args.only == "all":
        prec = fp64_headline(dev_x, s4, p4, st4, mask, run.out, ranks)

    full = args.only == "all"
    diff = bench_differentiable(args.diff_steps) if full and args.diff_steps > 0 else None
    batched = (bench_batched(args.batch_steps, autotune=not args.no_autotune)
               if full and args.batch_steps > 0 else None)
    noise_map = bench_noise_map(5) if full else None
    large = bench_large(args.large_steps) if full and args.large_steps > 0 else None
    public = bench_public_calls(2) if full else None

    split = None
    if args.only == "all" and args.precision == "auto":
        # the same workload with the split-BF16 three-product Gram
        pb = N.make_problem(s4, p4, st4, np.float32, N.MPPCA, "weighted", "bf16x3")
        pb.threshold_scale2 = 1.0
        pb.scale_f64 = 0
        pb.mask_kind = N.MASK_SPATIAL
        pb.mask_threshold = 50.0
        rb = N.DeviceRun(pb, dev_x, mask, None)
        rb.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rb.run()
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t0) / args.steps
        split = dict(precision="bf16x3", ms_per_step=tb * 1e3, value=vf / tb,
                     rel_l2_vs_fp32=_rel_l2(rb.out, run.out),
                     speedup_vs_fp32=(ms_per_step / 1e3) / tb)
        if prec is not None:
            split["rel_l2_vs_fp64"] = _rel_l2(rb.out, prec.pop("out64"))
        del rb
    if prec is not None:
        prec.pop("out64", None)

    plain = batched_t = None
    if not args.no_plain:
        pg = plain_gpu_baseline(data, sel, args.plain_stride)
        plain = dict(value=vf / pg["full_estimate_s"], unit="voxel-frames/s",
                     sample=(f"every {args.plain_stride}th selected patch "
                             f"({pg['sample_patches']} patches, {pg['seconds']:.1f} s), "
                             "extrapolated to all selected patches"),
                     per_patch_ms=pg["per_patch_ms"],
                     path="one patch per launch chain, no host synchronisation: torch gather, "
                          "Gram matmul, torch.linalg.eigh (FP32), MP-PCA rule, projection, "
                          "index_add_")
        bg = batched_gpu_baseline(data, sel, ranks)
        batched_t = dict(value=vf / bg["full_estimate_s"], unit="voxel-frames/s",
                         sample=(f"{bg['sample_patches']} of {len(sel)} selected patches in "
                                 f"batches of {bg['batch']} ({bg['seconds']:.1f} s), "
                                 "extrapolated to all selected patches"),
                         per_patch_ms=bg["per_patch_ms"],
                         rank_agreement_with_hip=bg["rank_agreement_with_hip"],
                         path="batched torch FP32: gather, bmm Gram, batched torch.linalg.eigh, "
                              "MP-PCA rule, projection, index_add_")

    line = {
        "metric": "MP-PCA voxel-frames/s at 80x80x48x300",
        "value": value,
        "unit": "voxel-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision != "fp64" else "f64",
        "data": "synthetic (seeded generator, patch_denoise.workloads.config2)",
        "config": {"workload": WORKLOAD, "shape": list(shape), "patch": 5, "overlap": 4,
                   "selected_patches": int(P), "mask": "ellipsoid (36,36,20)",
                   "precision": args.precision},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "stages": stages,
        "precision_vs_fp64": prec,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "speedup_vs_cpu_single_core": (value / cpu["single_core"]["value"]) if cpu else None,
        "split_bf16": split,
        "plain_gpu_baseline": plain,
        "speedup_vs_plain_gpu": (value / plain["value"]) if plain else None,
        "batched_torch_baseline": batched_t,
        "speedup_vs_batched_torch": (value / batched_t["value"]) if batched_t else None,
        "public_api_numpy": public,
        "differentiable": diff,
        "batched": batched,
        "noise_map": noise_map,
        "large_gram": large,
        "datagen_s": gen_s,
    }
    print(json.dumps(line, default=_jsonable))


def _jsonable(o):
    """0-d tensors / NumPy scalars in the result line become Python numbers."""
    if hasattr(o, "item"):
        print(f"bench: converted a {type(o).__name__} in the JSON line", file=sys.stderr)
        return o.item()
    raise TypeError(f"{type(o).__name__} is not JSON serialisable")


if __name__ == "__main__":
    main()
