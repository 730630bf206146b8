"""Differentiable optimal singular-value shrinkage (learnable threshold scales).

The optimal-shrinkage rule of the reference (space_time/lowrank.py:379-407)
computes, per patch, a noise level ``c = sqrt(T) * sigma`` and shrinks each
singular value ``s`` to ``c * f(s / c)``.  Here every patch carries a scale
``lambda_p`` with ``c_p = lambda_p * sqrt(T) * sigma_p``; ``lambda = 1``
reproduces the reference's output exactly.

``OptimalShrinkFunction`` is a ``torch.autograd.Function`` whose forward is
``pdamd_denoise`` and whose backward is ``pdamd_optimal_backward``: hand-written
HIP kernels that recompute each patch's SVD (Gram, Householder, all
eigenpairs) and apply the singular-vector adjoint with guarded
``(g_i - g_j) / (s_i - s_j)`` terms, the noise-level (median) dependence and
the centring and recombination adjoints.  Gradients flow to the volume and to
``lambda``; the retained rank (and so the weighted-recombination weight
``1 / (2 + r)``) is piecewise constant and contributes no gradient.
"""

import ctypes
import types

import numpy as np

from . import _native as N
from .space_time.base import spatial_variance_map
from .space_time.lowrank import OptimalSVDDenoiser
from .space_time.utils import marchenko_pastur_median

__all__ = ["OptimalShrinkage", "OptimalShrinkFunction", "optimal_shrinkage"]


class _Spec:
    """Problem description shared by the forward and backward calls."""

    def __init__(self, shape, dtype, patch_shape, patch_overlap, loss="fro",
                 recombination="weighted", mask=None, mask_threshold=50, noise_std=None,
                 precision="auto"):
        torch = N._torch()
        self.torch = torch
        den = OptimalSVDDenoiser(patch_shape, patch_overlap, loss=loss,
                                 recombination=recombination, precision=precision)
        meta = types.SimpleNamespace(shape=tuple(shape), dtype=np.dtype(dtype))
        prob, p_s, p_o, mask_arr = den._prepare(meta, mask, mask_threshold)
        beta = (p_s[-1] / np.prod(p_s[:-1]) if len(p_s) == len(shape)
                else shape[-1] / np.prod(p_s))
        den._set_rule(prob, marchenko_pastur_median(beta=beta), den.input_denoising_kwargs[
            "shrink_func"])
        self.template = bytes(prob)  # copied into a fresh struct per call
        dev = torch.device("cuda", torch.cuda.current_device())
        self.mask = None
        if mask_arr is not None:
            m = np.ascontiguousarray(mask_arr).astype(np.uint8).reshape(-1)
            self.mask = torch.from_numpy(m).to(dev)
        self.var = None
        if noise_std is not None:
            v = spatial_variance_map(noise_std, tuple(shape))
            self.var = torch.from_numpy(np.ascontiguousarray(v, np.float64).reshape(-1)).to(dev)
        self.shape = tuple(shape)
        self.shape4 = tuple(int(v) for v in prob.shape)
        grid = []
        for i in range(4):
            S, p, st = self.shape4[i], int(prob.patch[i]), int(prob.step[i])
            grid.append(((S - p) // st + 1) if (p < S and st > 0) else 1)
        self.grid4 = tuple(grid)
        self.npatch = int(np.prod(grid))
        self.selected = None
        # optional (ms, launches) ctypes arrays of length len(_native.STAGES):
        # per-stage kernel times accumulate there (HIP events on the stream)
        self.timers = None

    def problem(self, x, lam):
        prob = N.Problem.from_buffer_copy(self.template)
        prob.input_dev = x.data_ptr()
        prob.lam_dev = lam.data_ptr() if lam is not None else None
        prob.mask_dev = self.mask.data_ptr() if self.mask is not None else None
        prob.var_map_dev = self.var.data_ptr() if self.var is not None else None
        return prob

    def _ws(self, nbytes):
        return self.torch.empty(max(nbytes, 1), dtype=self.torch.uint8, device=self._dev())

    def _dev(self):
        return self.torch.device("cuda", self.torch.cuda.current_device())

    def _timers(self, prob):
        if self.timers is not None:
            prob.stage_ms = ctypes.cast(self.timers[0], ctypes.c_void_p).value
            prob.stage_launches = ctypes.cast(self.timers[1], ctypes.c_void_p).value

    def forward(self, x, lam):
        torch = self.torch
        prob = self.problem(x, lam)
        self._timers(prob)
        out = torch.empty(self.shape4, dtype=x.dtype, device=x.device)
        w = torch.empty(self.shape4, dtype=torch.float32, device=x.device)
        prob.output_dev = out.data_ptr()
        prob.weights_dev = w.data_ptr()
        nb = N.workspace_bytes(prob)
        ws = self._ws(nb)
        N._check(N.load_library().pdamd_denoise(ctypes.byref(prob), ws.data_ptr(), nb,
                                                torch.cuda.current_stream().cuda_stream))
        self.selected = int(prob.selected)
        return out, w

    def backward(self, x, lam, w, gout, want_lam=True):
        torch = self.torch
        prob = self.problem(x, lam)
        self._timers(prob)
        gx = torch.empty(self.shape4, dtype=x.dtype, device=x.device)
        glam = torch.empty(self.npatch, dtype=torch.float64, device=x.device) if want_lam else None
        nb = ctypes.c_size_t(0)
        lib = N.load_library()
        N._check(lib.pdamd_backward_workspace_size(ctypes.byref(prob), ctypes.byref(nb)))
        ws = self._ws(nb.value)
        N._check(lib.pdamd_optimal_backward(
            ctypes.byref(prob), gout.data_ptr(), w.data_ptr(), gx.data_ptr(),
            glam.data_ptr() if glam is not None else None, ws.data_ptr(), nb.value,
            torch.cuda.current_stream().cuda_stream))
        return gx, glam


def _function():
    torch = N._torch()

    class OptimalShrinkFunction(torch.autograd.Function):
        """out = optimal_shrinkage(x; lambda), with the hand-written adjoint."""

        @staticmethod
        def forward(ctx, x, lam, spec):
            xc = x.detach().reshape(spec.shape4).contiguous()
            lc = lam.detach().reshape(-1).to(torch.float64).contiguous()
            if lc.numel() != spec.npatch:
                raise ValueError(f"lambda has {lc.numel()} entries, the patch grid "
                                 f"{spec.grid4} has {spec.npatch}")
            out, w = spec.forward(xc, lc)
            ctx.spec = spec
            ctx.lam_shape, ctx.lam_dtype = lam.shape, lam.dtype
            ctx.save_for_backward(xc, lc, w)
            return out.reshape(x.shape)

        @staticmethod
        def backward(ctx, gout):
            xc, lc, w = ctx.saved_tensors
            spec = ctx.spec
            g = gout.reshape(spec.shape4).to(xc.dtype).contiguous()
            gx, glam = spec.backward(xc, lc, w, g, want_lam=ctx.needs_input_grad[1])
            gx = gx.reshape(gout.shape) if ctx.needs_input_grad[0] else None
            if glam is not None:
                glam = glam.reshape(ctx.lam_shape).to(ctx.lam_dtype)
            return gx, glam, None

    return OptimalShrinkFunction


_FN = None


def __getattr__(name):
    if name == "OptimalShrinkFunction":
        return _get_fn()
    raise AttributeError(name)


def _get_fn():
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN


def optimal_shrinkage(x, lam, patch_shape, patch_overlap, loss="fro", recombination="weighted",
                      mask=None, mask_threshold=50, noise_std=None, precision="auto", spec=None):
    """Differentiable optimal thresholding of a CUDA tensor ``x`` (last axis =
    time) with per-patch threshold scales ``lam`` (one per patch of the grid,
    any shape with that many entries).  Returns the denoised tensor; the
    output matches ``optimal_thresholding(x, ...)[0]`` when ``lam`` is all ones.
    """
    if spec is None:
        spec = _Spec(tuple(x.shape), _np_dtype(x), patch_shape, patch_overlap, loss,
                     recombination, mask, mask_threshold, noise_std, precision)
    return _get_fn().apply(x, lam, spec)


def _np_dtype(x):
    return np.dtype(str(x.dtype).replace("torch.", ""))


class OptimalShrinkage:
    """Optimal-shrinkage denoiser with a learnable per-patch threshold scale.

    ``OptimalShrinkage(shape, dtype, patch_shape, patch_overlap, ...)`` holds
    ``lam`` (a ``torch.nn.Parameter`` shaped like the patch grid, initialised
    to 1) and applies :func:`optimal_shrinkage` in ``__call__``.
    """

    def __init__(self, shape, dtype, patch_shape, patch_overlap, loss="fro",
                 recombination="weighted", mask=None, mask_threshold=50, noise_std=None,
                 precision="auto", lam_init=1.0):
        torch = N._torch()
        self.spec = _Spec(tuple(shape), dtype, patch_shape, patch_overlap, loss, recombination,
                          mask, mask_threshold, noise_std, precision)
        grid = tuple(g for g in self.spec.grid4)
        self.lam = torch.nn.Parameter(
            torch.full(grid, float(lam_init), dtype=torch.float32, device="cuda"))

    def parameters(self):
        return [self.lam]

    def __call__(self, x):
        return _get_fn().apply(x, self.lam, self.spec)
