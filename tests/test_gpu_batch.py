"""Batched multi-volume launch (pdamd_denoise_batch) against per-volume calls.

Each volume of the batch must come out as the same functional-API call made
alone (those are pinned to the reference fixtures elsewhere); the captured
graph must replay to the same results, with eager work in between.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.nan_to_num(np.asarray(a, np.float64)), np.nan_to_num(np.asarray(b, np.float64))
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _calls(scale=0.3):
    from patch_denoise import workloads

    vols = workloads.config4(scale=scale)
    calls, skipped = workloads.config4_calls(vols)
    return calls, skipped


def _single(calls):
    import patch_denoise as pd

    return [getattr(pd, name)(*args, **kw) for name, args, kw in calls]


def test_batch_matches_single_calls():
    from patch_denoise.batch import denoise_batch

    calls, skipped = _calls()
    print(f"[batch] {len(calls)} volumes, skipped {skipped}", flush=True)
    assert len(calls) == 8 and not skipped  # includes the 7^3 x 200 volume (n = 200)
    ref = _single(calls)
    out = denoise_batch(calls, use_graph=False)
    for k, ((o, w, _, _), (ro, rw, _, _)) in enumerate(zip(out, ref)):
        e, ew = rel_l2(o, ro), rel_l2(w, rw)
        print(f"[batch] volume {k} ({calls[k][0]}): rel-L2 out {e:.2e} weights {ew:.2e}", flush=True)
        assert e < 1e-6 and ew == 0.0


def test_batch_graph_replay():
    import torch

    from patch_denoise.batch import BatchPlan

    calls, _ = _calls()
    ref = _single(calls)
    plan = BatchPlan(calls, use_graph=True)
    for rep in range(3):
        res = plan.run().results()
        _single(calls[:1])  # eager work between replays
        torch.cuda.synchronize()
        for k, ((o, w, _, _), (ro, rw, _, _)) in enumerate(zip(res, ref)):
            e = rel_l2(o, ro)
            assert e < 1e-6 and rel_l2(w, rw) == 0.0, (rep, k, e)
        print(f"[batch] replay {rep}: ok", flush=True)


def test_autotuned_batch_matches_single_calls():
    """Launch shapes chosen by patch_denoise.autotune: same results (only the
    atomic summation order of overlapping patches may differ)."""
    from patch_denoise import autotune
    from patch_denoise.batch import BatchPlan

    calls, _ = _calls()
    ref = _single(calls)
    autotune.clear_cache()
    plan = BatchPlan(calls, use_graph=True, autotune=True)
    print(f"[batch] autotuned launch_per_cu: {plan.tuned}", flush=True)
    assert all(all(0 <= v <= 32 for v in t) for t in plan.tuned)  # 0: library default
    for rep in range(2):
        res = plan.run().results()
        for k, ((o, w, _, _), (ro, rw, _, _)) in enumerate(zip(res, ref)):
            e, ew = rel_l2(o, ro), rel_l2(w, rw)
            assert e < 1e-6 and ew < 1e-7, (rep, k, e, ew)
    print("[batch] autotuned replays: ok", flush=True)


def test_batch_reports_unsupported_volume():
    from patch_denoise import _native as N
    from patch_denoise.batch import denoise_batch

    rng = np.random.default_rng(0)
    ok = rng.standard_normal((8, 8, 8, 30)).astype(np.float32) + 10
    # n = min(13^3, 2200) = 2197: above the largest supported Gram (2048)
    big = rng.standard_normal((13, 13, 13, 2200)).astype(np.float32) + 10
    with pytest.raises(N.NativeError, match="volume 1"):
        denoise_batch([("mp_pca", (ok, 3, 2), {}), ("mp_pca", (big, 13, 12), {})])
