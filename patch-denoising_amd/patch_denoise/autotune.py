"""Launch-shape autotuner for ``pdamd_denoise`` (no counterpart in the reference).

The library launches its per-patch kernels with a default number of
workgroups per CU for each stage (Gram, Householder, eigen, synthesis), capped
by the kernel's LDS.  The best value depends on the patch shape, the Gram
size, the dtype, the rule and the patch count, so :func:`tune` measures it on
the caller's own problem: one stage at a time (coordinate descent, the other
stages at their current best), each candidate timed with the library's
per-stage HIP-event timers over one full call; the library's default (0) is a
candidate too and is replaced only by a value at least ``margin`` (5%) faster
on that stage, so tuning never trades a measured default for noise.  Winners are
cached by problem shape for the life of the process and written into
``prob.launch_per_cu``; a batch (``patch_denoise.batch.BatchPlan(...,
autotune=True)``) tunes every volume before its graph is captured.

Only launch grids change: every patch is computed the same way, so outputs
differ at most by the order of the float atomic additions of the overlapping
patches.
"""

import ctypes

from . import _native as N

# stage -> (index in pdamd_problem.launch_per_cu, candidate workgroups per CU)
CANDIDATES = {
    "gram": (0, (2, 4, 8)),
    "tridiag": (1, (2, 4, 8)),
    "trieig": (2, (4, 8, 16, 32)),
    "synth": (3, (2, 4, 8)),
}

_CACHE = {}


def problem_key(prob):
    """Shape, dtype, rule and size class of a problem (the cache key)."""
    npatch = 1
    for i in range(4):
        S, p, st = int(prob.shape[i]), int(prob.patch[i]), int(prob.step[i])
        npatch *= ((S - p) // st + 1) if (p < S and st > 0) else 1
    return (int(prob.dtype), tuple(int(v) for v in prob.patch), int(prob.method),
            int(prob.precision), int(prob.recombination), max(npatch, 1).bit_length())


def _stage_ms(run, prob):
    ms = (ctypes.c_float * len(N.STAGES))()
    cnt = (ctypes.c_int64 * len(N.STAGES))()
    prob.stage_ms = ctypes.cast(ms, ctypes.c_void_p).value
    prob.stage_launches = ctypes.cast(cnt, ctypes.c_void_p).value
    try:
        run.run()
    finally:
        prob.stage_ms = None
        prob.stage_launches = None
    return {name: ms[i] for i, name in enumerate(N.STAGES)}


def tune(prob, data, mask=None, var_map=None, candidates=None, verbose=False, margin=0.05):
    """Choose ``prob.launch_per_cu`` for this problem (cached); returns the
    chosen tuple and, when measured now, the per-candidate stage times."""
    key = problem_key(prob)
    if key in _CACHE:
        best = _CACHE[key]
        for i, v in enumerate(best):
            prob.launch_per_cu[i] = v
        return best, None
    cands = candidates or CANDIDATES
    for i in range(4):
        prob.launch_per_cu[i] = 0
    run = N.DeviceRun(prob, data, mask, var_map)
    run.run()  # warm-up (first launches, code loading)
    report = {}
    for stage, (slot, values) in cands.items():
        times = {}
        for v in (0,) + tuple(values):
            prob.launch_per_cu[slot] = v
            times[v] = _stage_ms(run, prob)[stage]
        best_v = min(times, key=times.get)
        # keep the default unless clearly beaten (also when the stage did not run)
        if not times[0] > 0 or times[best_v] > (1.0 - margin) * times[0]:
            best_v = 0
        prob.launch_per_cu[slot] = best_v
        report[stage] = times
        if verbose:
            print(f"autotune {stage}: " + ", ".join(f"{v}/CU {t:.2f} ms" for v, t in times.items())
                  + f" -> {best_v}", flush=True)
    del run
    best = tuple(int(prob.launch_per_cu[i]) for i in range(4))
    _CACHE[key] = best
    return best, report


def clear_cache():
    _CACHE.clear()
