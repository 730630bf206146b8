"""Batched denoising of several volumes in one device call.

The reference denoises one volume per call (``denoise.py``); a list of calls
is a Python loop over ``BaseSpaceTimeDenoiser.denoise`` (space_time/base.py:
177-270).  ``denoise_batch`` takes that list of calls, runs each call's host
preamble exactly as the single call would (patch parameters, NORDIC's
Monte-Carlo threshold from the global ``np.random`` stream in call order, the
Marchenko-Pastur median, prior variance maps), then denoises every volume in
ONE ``pdamd_denoise_batch`` call: selection for all volumes, one host
synchronisation for all patch counts, then each volume's kernel chain on its
own forked stream (small volumes run concurrently), optionally replayed from
a captured hipGraph.

    outs = denoise_batch([
        ("mp_pca", (vol_a, 5, 4), {"mask": mask_a}),
        ("raw_svt", (vol_b, 3, 2), {"threshold": 40.0}),
    ])
    denoised_a, weights_a, noise_a, rank_a = outs[0]
"""

from . import _native as N
from . import denoise as _functional
from .space_time import base as _base

__all__ = ["denoise_batch", "BatchPlan"]

_FUNCS = ("mp_pca", "hybrid_pca", "raw_svt", "nordic", "optimal_thresholding",
          "adaptive_thresholding")


def _collect(calls):
    jobs = []
    _base._COLLECT.append(jobs)
    try:
        for name, args, kwargs in calls:
            if name not in _FUNCS:
                raise ValueError(f"unknown entry point {name!r}; expected one of {_FUNCS}")
            getattr(_functional, name)(*args, **(kwargs or {}))
    finally:
        _base._COLLECT.pop()
    if len(jobs) != len(calls):
        raise RuntimeError("a call did not reach the device stage")
    return jobs


class BatchPlan:
    """Prepared batch: device buffers allocated once, reusable with
    :meth:`run` (the second and later runs replay the captured graph when
    ``use_graph``).  Inputs stay as given at construction.  ``autotune``
    chooses each volume's launch shapes first (``patch_denoise.autotune``)."""

    def __init__(self, calls, use_graph=True, autotune=False, want_patch_info=False):
        self.jobs = _collect(list(calls))
        self.tuned = None
        if autotune:  # launch shapes per volume, before any graph is captured
            from . import autotune as _tune

            self.tuned = [_tune.tune(j.prob, j.data, j.mask_arr, j.var_map)[0] for j in self.jobs]
        self.batch = N.BatchRun([(j.prob, j.data, j.mask_arr, j.var_map) for j in self.jobs],
                                use_graph=use_graph, want_patch_info=want_patch_info)

    def run(self):
        self.batch.run()
        return self

    def patch_info(self):
        """Per volume: selected patch indices, retained ranks, noise and top
        eigenvalues (``want_patch_info=True`` at construction)."""
        return self.batch.patch_info()

    def results(self):
        res = self.batch.results([j.data.shape for j in self.jobs])
        return [j.finish(out, w, n) for j, (out, w), n in zip(self.jobs, res, self.batch.selected)]


def denoise_batch(calls, use_graph=False):
    """Run a list of ``(entry_point_name, args, kwargs)`` calls of the
    functional API as one batched device call; returns the list of their
    4-tuples, in order."""
    return BatchPlan(calls, use_graph=use_graph).run().results()
