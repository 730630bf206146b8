"""Run the CPU oracle over many patches in parallel.

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``: used by the tests to
check whole volumes and by the benchmark's CPU-baseline leg to time the
oracle on one process per host core.

The GPU tests denoise whole volumes; checking them against the oracle patch by
patch takes minutes on one core, so the work is spread over a pool of worker
processes started with ``spawn`` (a fork of a process that has initialised the
GPU is not safe).  Volumes reach the workers as a read-only ``np.memmap`` of a
temporary file, never through pickling.  Each worker runs one BLAS thread.

Per-method patch rules (``method``, ``params``):

* ``"mppca"``   ``{"ts": threshold_scale}``          oracle ``mppca_patch``
* ``"hybrid"``  ``{"var": spatial variance map}``    oracle ``hybrid_patch``
* ``"optimal"`` ``{"loss", "mpm", "var" or None}``   oracle ``optimal_patch``
* ``"raw"``     ``{"tau": threshold}``               oracle ``raw_patch``

Besides the rank, every patch reports its *decision margin*: the smallest
distance, over every comparison the rule makes, between the two compared
quantities, divided by the patch's largest eigenvalue (eigen rules) or largest
singular value (SVD rules).  FP32 eigensolvers (LAPACK's and the GPU's) agree
only to a few units of float32 rounding of that largest value, so a rank can
legitimately differ only where this margin is that small: a tie at the cutoff.
"""

import os
import tempfile
from concurrent.futures import ProcessPoolExecutor
import multiprocessing as mp

import numpy as np

N_WORKERS = max(1, min(16, os.cpu_count() or 1))


class SharedVolume:
    """A volume written once to a temporary file; workers memory-map it."""

    def __init__(self, data):
        fd, self.path = tempfile.mkstemp(suffix=".vol")
        os.close(fd)
        mm = np.memmap(self.path, dtype=data.dtype, mode="w+", shape=data.shape)
        mm[...] = data
        mm.flush()
        del mm
        self.shape, self.dtype = data.shape, data.dtype.str

    def spec(self):
        return (self.path, self.shape, self.dtype)

    def close(self):
        if self.path and os.path.exists(self.path):
            os.unlink(self.path)
        self.path = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _open(spec):
    path, shape, dtype = spec
    return np.memmap(path, dtype=np.dtype(dtype), mode="r", shape=tuple(shape))


def _init_worker():
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(limits=1)
    except ImportError:
        pass


# ---------------------------------------------------------------------------
# per-patch rule + decision margin
# ---------------------------------------------------------------------------


def _mppca_margin(a, M, T, ts):
    """Smallest gap of the MP-PCA rule's comparisons (lowrank.py:48-54), as
    the rule is evaluated: the loop tests and the final count."""
    gaps = []
    k = 0
    var = np.mean(a) * 4 * np.sqrt((len(a) - k + 1) / M)
    while k < len(a):
        gaps.append(abs(float(var) - float(a[~k] - a[0])))
        if not var < a[~k] - a[0]:
            break
        k += 1
        var = np.mean(a[:-k]) * 4 * np.sqrt((T - k + 1) / M)
    if k < len(a):
        noise = np.mean(a[: len(a) - k])
        cut = noise * ts**2
        gaps.extend(np.abs(a.astype(np.float64) - float(cut)).tolist())
    return min(gaps) / max(abs(float(a[-1])), 1e-300)


def _hybrid_margin(a, varest):
    gaps = []
    k = 0
    var = np.mean(a)
    while k < len(a) - 2:
        gaps.append(abs(float(var) - float(varest)))
        if not var > varest:
            break
        k += 1
        var = np.mean(a[:-k])
    return min(gaps) / max(abs(float(a[-1])), 1e-300) if gaps else np.inf


def _run_patch(O, method, params, X, L, prior):
    """(patch, rank, margin) for the oracle rule ``method`` on Casorati X."""
    if method == "mppca":
        P, r, _ = O.mppca_patch(X, params["ts"])
        _, a, _, _ = O.eig_analysis(X)
        return P, r, _mppca_margin(a, X.shape[0], X.shape[1], params["ts"])
    if method == "hybrid":
        varest = np.mean(prior.get(L))
        P, r, _ = O.hybrid_patch(X, varest)
        _, a, _, _ = O.eig_analysis(X)
        return P, r, _hybrid_margin(a, varest)
    if method == "raw":
        P, r, _ = O.raw_patch(X, params["tau"])
        s = np.linalg.svd(X - X.mean(axis=0), compute_uv=False)
        return P, r, float(np.min(np.abs(s - params["tau"])) / max(s[0], 1e-300))
    if method == "optimal":
        shrink = O.SHRINKERS[params["loss"]]
        vp = None if prior is None else prior.get(L)
        P, r, _ = O.optimal_patch(X, shrink, params["mpm"], vp)
        s = np.linalg.svd(X - X.mean(axis=0), compute_uv=False)
        T = X.shape[1]
        sigma = np.mean(np.sqrt(vp)) if vp is not None else np.median(s) / np.sqrt(T * params["mpm"])
        c = np.sqrt(T) * sigma
        edge = 1 + np.sqrt(T / X.shape[0])
        return P, r, float(np.min(np.abs(s / c - edge)) * c / max(s[0], 1e-300))
    raise ValueError(method)


def _prior(method, params, shape, ps, po):
    from oracle import cpu_ref as O

    var = params.get("var") if method in ("hybrid", "optimal") else None
    if var is None:
        return None
    return O.StridedPatches(np.broadcast_to(var[..., None], shape), ps, po)


def _ranks_task(spec, ps, po, method, params, indices):
    from oracle import cpu_ref as O

    data = _open(spec)
    src = O.StridedPatches(data, ps, po)
    prior = _prior(method, params, data.shape, ps, po)
    M = int(np.prod(ps[:-1]))
    ranks = np.empty(len(indices), np.int64)
    margins = np.empty(len(indices), np.float64)
    for j, L in enumerate(indices):
        X = np.array(src.get(int(L))).reshape(M, -1)
        _, ranks[j], margins[j] = _run_patch(O, method, params, X, int(L), prior)
    return ranks, margins


def _window_task(spec, ps, po, method, params, indices, lo, hi):
    """Weighted recombination (theta = 1/(2+r)) of the given patches,
    restricted to the spatial window lo..hi, in float64."""
    from oracle import cpu_ref as O

    data = _open(spec)
    src = O.StridedPatches(data, ps, po)
    prior = _prior(method, params, data.shape, ps, po)
    step = [p - o for p, o in zip(ps, po)]
    ws = tuple(h - l for l, h in zip(lo, hi))
    T = data.shape[-1]
    out = np.zeros(ws + (T,), np.complex128 if np.iscomplexobj(data) else np.float64)
    w = np.zeros(ws + (T,), np.float64)
    M = int(np.prod(ps[:-1]))
    ranks = {}
    for L in indices:
        g = np.unravel_index(int(L), src.grid)
        o = [gi * st for gi, st in zip(g[:3], step[:3])]
        X = np.array(src.get(int(L))).reshape(M, -1)
        P, r, _ = _run_patch(O, method, params, X, int(L), prior)
        P = P.reshape(ps)
        ranks[int(L)] = int(r)
        th = 1.0 / (2 + r)
        a = [max(lo[d], o[d]) for d in range(3)]
        b = [min(hi[d], o[d] + ps[d]) for d in range(3)]
        so = tuple(slice(a[d] - lo[d], b[d] - lo[d]) for d in range(3))
        sp = tuple(slice(a[d] - o[d], b[d] - o[d]) for d in range(3))
        out[so] += th * P[sp]
        w[so] += th
    return out, w, ranks


# ---------------------------------------------------------------------------
# parallel drivers
# ---------------------------------------------------------------------------


def _pool(workers=None):
    return ProcessPoolExecutor(max_workers=workers or N_WORKERS,
                               mp_context=mp.get_context("spawn"), initializer=_init_worker)


def _warm_task(_):
    from oracle import cpu_ref as O

    rng = np.random.default_rng(0)
    O.mppca_patch(rng.standard_normal((27, 40)).astype(np.float32), 1.0)
    return os.getpid()


def _mppca_time_task(spec, ps, po, ts, indices):
    """MP-PCA of the given patches with weighted accumulation into a
    worker-local volume, as the reference's patch loop does
    (space_time/base.py:96-110, lowrank.py:24-63); returns the seconds spent."""
    import time

    from oracle import cpu_ref as O

    data = _open(spec)
    src = O.StridedPatches(data, ps, po)
    out = np.zeros(data.shape, np.float32)
    w = np.zeros(data.shape, np.float32)
    dst, wts = O.StridedPatches(out, ps, po), O.StridedPatches(w, ps, po)
    M = int(np.prod(ps[:-1]))
    t0 = time.perf_counter()
    for L in indices:
        X = np.array(src.get(int(L))).reshape(M, -1)
        P, r, _ = O.mppca_patch(X, ts)
        th = 1 / (2 + r)
        o = dst.get(int(L))
        o += P.reshape(ps) * th
        ww = wts.get(int(L))
        ww += th
    return time.perf_counter() - t0


def time_mppca_pool(vol, ps, po, ts, indices, workers=None):
    """Wall time of MP-PCA over ``indices`` on a pool of ``workers`` processes
    (one BLAS thread each), the pool started and warmed beforehand.  The
    final sum of the workers' partial volumes is not included."""
    import time

    workers = workers or N_WORKERS
    indices = np.asarray(indices, np.int64)
    parts = np.array_split(indices, max(1, min(len(indices), 4 * workers)))
    pool = _pool(workers)
    try:
        list(pool.map(_warm_task, range(4 * workers)))
        t0 = time.perf_counter()
        futs = [pool.submit(_mppca_time_task, vol.spec(), ps, po, ts, p) for p in parts
                if len(p)]
        busy = sum(f.result() for f in futs)
        wall = time.perf_counter() - t0
    finally:
        pool.shutdown()
    return dict(wall_s=wall, busy_s=busy, workers=workers, patches=int(len(indices)))


def oracle_ranks(vol, ps, po, method, params, indices, pool=None):
    """Oracle rank and decision margin of every patch in ``indices``."""
    indices = np.asarray(indices, np.int64)
    parts = np.array_split(indices, max(1, min(len(indices), 8 * N_WORKERS)))
    own = pool is None
    pool = pool or _pool()
    try:
        futs = [pool.submit(_ranks_task, vol.spec(), ps, po, method, params, p) for p in parts
                if len(p)]
        res = [f.result() for f in futs]
    finally:
        if own:
            pool.shutdown()
    return np.concatenate([r for r, _ in res]), np.concatenate([m for _, m in res])


def covering(sel, grid, ps, po, lo, hi):
    """Selected patches whose spatial footprint meets the window lo..hi."""
    step = np.array([p - o for p, o in zip(ps, po)][:3])
    g = np.array(np.unravel_index(np.asarray(sel, np.int64), grid)[:3]).T
    o = g * step
    p = np.array(ps[:3])
    keep = np.all((o + p > np.array(lo)) & (o < np.array(hi)), axis=1)
    return np.asarray(sel)[keep]


def window_weights32(ps, po, grid, ranks, lo, hi, T):
    """Weights of the window lo..hi summed as the reference sums them: a
    float32 array, each patch (increasing linear index) adding its theta =
    1/(2+r) as f32(f64(w) + theta) (space_time/base.py:243-256, with NumPy's
    float64 scalar theta); ``ranks`` maps every patch meeting the window to
    its rank (``oracle_window``'s third result)."""
    step = [p - o for p, o in zip(ps, po)]
    ws = tuple(h - l for l, h in zip(lo, hi))
    w = np.zeros(ws + (T,), np.float32)
    for L in sorted(ranks):
        g = np.unravel_index(int(L), grid)
        o = [gi * st for gi, st in zip(g[:3], step[:3])]
        a = [max(lo[d], o[d]) for d in range(3)]
        b = [min(hi[d], o[d] + ps[d]) for d in range(3)]
        so = tuple(slice(a[d] - lo[d], b[d] - lo[d]) for d in range(3))
        th = np.float64(1.0) / np.float64(2 + ranks[L])
        w[so] = (w[so].astype(np.float64) + th).astype(np.float32)
    return w


def oracle_window(vol, ps, po, method, params, sel, grid, lo, hi, pool=None):
    """Oracle output / weights on the window and the ranks of its patches."""
    idx = covering(sel, grid, ps, po, lo, hi)
    parts = np.array_split(idx, max(1, min(len(idx), 2 * N_WORKERS)))
    own = pool is None
    pool = pool or _pool()
    try:
        futs = [pool.submit(_window_task, vol.spec(), ps, po, method, params, p, lo, hi)
                for p in parts if len(p)]
        res = [f.result() for f in futs]
    finally:
        if own:
            pool.shutdown()
    out = sum(r[0] for r in res)
    w = sum(r[1] for r in res)
    ranks = {}
    for r in res:
        ranks.update(r[2])
    with np.errstate(invalid="ignore", divide="ignore"):
        return out / w, w, ranks
