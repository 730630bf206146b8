"""Patch geometry and the GPU patch driver (mirrors space_time/base.py).

``BaseSpaceTimeDenoiser.denoise`` (reference: space_time/base.py:177-270)
hands the whole volume to the MI355X library in one call: selection, the
per-patch low-rank step of the subclass's rule, recombination and
normalisation all run on the GPU.  ``PatchedArray`` (base.py:13-150) is kept
as the same host-side strided-view helper for callers that use it.
"""

import abc
import logging

import numpy as np

from .. import _native as N


class PatchedArray:
    """Sliding patch view of an array (reference: space_time/base.py:13-150).

    Grid per dimension ``(S - p) // (p - o) + 1`` when ``p < S`` else 1, patch
    clipped to the array, no padding.  ``ValueError`` when an overlap exceeds
    the patch or the number of dimensions does not match.
    """

    def __init__(self, array, patch_shape, patch_overlap, dtype=None, padding_mode="edge",
                 **kwargs):
        if isinstance(array, tuple):
            array = np.zeros(array, dtype=dtype)
        self._arr = array
        self._ps = np.asarray(patch_shape)
        self._po = patch_overlap
        step = self._ps - np.asarray(self._po)
        if np.any(step < 0):
            raise ValueError("overlap should be smaller than patch on every dimension.")
        if self._ps.size != array.ndim or step.size != array.ndim:
            raise ValueError(
                "self._ps and step must have the same number of dimensions as the "
                "input self._array."
            )
        self._ps = np.minimum(self._ps, array.shape)
        grid = _grid(array.shape, self._ps, step)
        strides = tuple(
            array.strides[i] * step[i] if self._ps[i] < array.shape[i] else 0
            for i in range(array.ndim)
        ) + array.strides
        self.sliding_view = np.lib.stride_tricks.as_strided(
            array, shape=grid + tuple(self._ps), strides=strides
        )
        self._grid_shape = grid

    @property
    def n_patches(self):
        return int(np.prod(self._grid_shape))

    def get_patch(self, idx):
        return self.sliding_view[np.unravel_index(idx, self._grid_shape)]

    def set_patch(self, idx, value):
        """No-op, as in the reference (base.py:93-95)."""
        self.sliding_view[np.unravel_index(idx, self._grid_shape)]

    def idx2slice(self, idx):
        g = np.unravel_index(idx, self._grid_shape)
        return tuple(slice(i * (p - o), i * (p - o) + p) for i, p, o in zip(g, self._ps, self._po))

    @classmethod
    def linear_to_patch_indices(cls, idx, data_shape, patch_shape, patch_overlap):
        step = np.asarray(patch_shape) - np.asarray(patch_overlap)
        g = np.unravel_index(idx, _grid(data_shape, patch_shape, step))
        return tuple(i * s for i, s in zip(g, step))

    def add2patch(self, idx, value):
        patch = self.get_patch(idx)
        patch += value

    def __getattr__(self, name):
        return getattr(self._arr, name)


def _grid(shape, ps, step):
    return tuple(
        int((shape[i] - ps[i]) // step[i] + 1) if (ps[i] < shape[i] and step[i] > 0) else 1
        for i in range(len(shape))
    )


def _geometry(data_shape, p_s, p_o):
    """Validate like ``PatchedArray`` and pad to the library's 4-D layout.

    Returns ``(shape4, patch4, step4)`` with the spatial dims padded by leading
    ones.  Raises the reference's ``ValueError``s for a negative step or a rank
    mismatch.
    """
    ndim = len(data_shape)
    ps = np.asarray(p_s)
    step = ps - np.asarray(p_o)
    if np.any(step < 0):
        raise ValueError("overlap should be smaller than patch on every dimension.")
    if ps.size != ndim or step.size != ndim:
        raise ValueError(
            "self._ps and step must have the same number of dimensions as the input self._array."
        )
    if ndim < 2 or ndim > 4:
        raise ValueError(f"input must have 1 to 3 spatial dimensions plus time, got {ndim - 1}")
    clipped = np.minimum(ps, data_shape)
    if int(np.prod(clipped[:-1])) != int(np.prod(ps[:-1])) or clipped[-1] != ps[-1]:
        raise ValueError(
            f"patch {tuple(int(v) for v in ps)} is larger than the input {data_shape}; "
            "the Casorati matrix would not be well formed"
        )
    pad = 4 - ndim
    shape4 = (1,) * pad + tuple(int(v) for v in data_shape)
    patch4 = (1,) * pad + tuple(int(v) for v in clipped)
    step4 = (1,) * pad + tuple(int(v) for v in step)
    return shape4, patch4, step4


# Stack of job lists: while non-empty, denoise() records its prepared problem
# in the innermost list instead of running it (patch_denoise.batch).
_COLLECT = []


class _Deferred:
    """A configured denoise() call waiting for a batched run."""

    def __init__(self, denoiser, prob, data, mask_arr, var_map, p_s, p_o, progbar):
        self.denoiser, self.prob, self.data = denoiser, prob, data
        self.mask_arr, self.var_map, self.p_s, self.p_o = mask_arr, var_map, p_s, p_o
        self.progbar = progbar

    def finish(self, out, w, n_selected):
        """The 4-tuple denoise() would have returned."""
        self.denoiser.n_selected = n_selected
        _progress(self.progbar, n_selected)
        shape = self.data.shape
        return (out, w, PatchedArray(shape, self.p_s, self.p_o, dtype=np.float32),
                PatchedArray(shape, self.p_s, self.p_o, dtype=np.int32))


class BaseSpaceTimeDenoiser(abc.ABC):
    """Patch-based low-rank denoiser for space x time data (base.py:154-384).

    Subclasses set ``_rule`` (the threshold rule run on the GPU) and
    ``_configure(prob, input_data, p_s)`` (its scalar operands).

    Extra keyword ``precision`` ("auto", "fp32", "fp64") selects the eigen
    stage's arithmetic; "auto" uses FP32 for the eigenvalue rules on float32
    data and FP64 for the singular-value rules.
    """

    _rule = None

    def __init__(self, patch_shape, patch_overlap, recombination="weighted", precision="auto"):
        self.p_shape = patch_shape
        self.p_ovl = patch_overlap
        if recombination not in ["weighted", "average", "center"]:
            raise ValueError("recombination must be one of 'weighted', 'average', 'center'")
        self.recombination = recombination
        if precision not in N.PREC:
            raise ValueError(f"precision must be one of {sorted(N.PREC)}")
        self.precision = precision
        self.input_denoising_kwargs = dict()
        self.n_selected = None

    # ---- host preamble shared by every rule -------------------------------------
    def _prepare(self, input_data, mask, mask_threshold):
        data_shape = input_data.shape
        p_s, p_o = self._get_patch_param(data_shape)
        shape4, patch4, step4 = _geometry(data_shape, p_s, p_o)
        if mask is None:
            mask_arr, kind = None, N.MASK_NONE
        elif mask.shape == data_shape:
            mask_arr, kind = mask, N.MASK_FULL
        elif mask.shape == data_shape[:-1]:
            mask_arr, kind = mask, N.MASK_SPATIAL
        else:
            raise ValueError(f"Mask shape {mask.shape} is incompatible with input {data_shape}.")
        prob = N.make_problem(shape4, patch4, step4, input_data.dtype, self._rule,
                              self.recombination, self.precision)
        prob.mask_kind = kind
        prob.mask_threshold = float(mask_threshold)
        return prob, p_s, p_o, mask_arr

    def denoise(self, input_data, mask=None, mask_threshold=50, progbar=None):
        """Denoise ``input_data`` (last axis = time); see the reference's
        ``BaseSpaceTimeDenoiser.denoise`` (base.py:177-270).

        Returns ``(denoised, weights, noise_std_estimate, rank_map)``; the last
        two are all-zero ``PatchedArray`` maps, as in the reference.
        """
        input_data = np.asarray(input_data)
        N.dtype_code(input_data.dtype)
        prob, p_s, p_o, mask_arr = self._prepare(input_data, mask, mask_threshold)
        var_map = self._configure(prob, input_data, p_s)
        if _COLLECT:  # inside patch_denoise.batch.denoise_batch: run later, batched
            job = _Deferred(self, prob, input_data, mask_arr, var_map, p_s, p_o, progbar)
            _COLLECT[-1].append(job)
            return job
        run = N.DeviceRun(prob, input_data, mask_arr, var_map).run()
        out, w = run.result(input_data.shape)
        self.n_selected = run.selected
        del run
        _progress(progbar, self.n_selected)
        shape = input_data.shape
        noise = PatchedArray(shape, p_s, p_o, dtype=np.float32)
        rank = PatchedArray(shape, p_s, p_o, dtype=np.int32)
        return out, w, noise, rank

    def _configure(self, prob, input_data, p_s):
        """Fill rule operands into ``prob``; return a spatial variance map or None."""
        return None

    # ---- the plugin method, on the GPU for a batch of one ------------------------
    @abc.abstractmethod
    def _patch_processing(self, patch, patch_idx=None, **kwargs):
        """Process one Casorati matrix (base.py:352-357)."""

    def _process_casorati(self, patches, var_rows=None, configure=None):
        """Run the rule on a batch of Casorati matrices ``(B, M, T)`` on the GPU.

        Returns ``(denoised (B, M, T), ranks (B,), noise (B,), top10 (B, 10))``.
        """
        patches = np.asarray(patches)
        B, M, T = patches.shape
        prob = N.make_problem((B, M, 1, T), (1, M, 1, T), (1, 1, 1, 1), patches.dtype,
                              self._rule, "average", self.precision)
        (configure or self._configure_patch)(prob, patches)
        run = N.DeviceRun(prob, patches, None, var_rows, want_patch_info=True).run()
        out, _ = run.result(patches.shape)
        info = run.patch_info()
        return out, info["rank"], info["noise"], info["eig"]

    def _configure_patch(self, prob, patches):
        raise NotImplementedError

    def _get_patch_param(self, data_shape):
        """Per-dimension patch shape and overlap (base.py:359-384).

        Both specs go through the same expansion; the ill-conditioning
        warning (same text as the reference) fires when the patch has fewer
        voxels than the volume has frames.
        """
        shape = _expand_patch_spec(self.p_shape, data_shape)
        overlap = _expand_patch_spec(self.p_ovl, data_shape)
        if np.prod(shape[:-1]) < data_shape[-1]:
            logging.warning(
                f"the number of voxel in patch ({np.prod(shape)}) is smaller than the"
                f" last dimension ({data_shape[-1]}), this makes an ill-conditioned"
                "matrix for SVD.",
                stacklevel=2,
            )
        return shape, overlap


def _expand_patch_spec(spec, data_shape):
    """Patch spec as a per-axis tuple (base.py:365-373).

    An integer applies to every spatial axis; a list becomes a tuple.  A spec
    naming only the spatial axes is completed with the full time length (one
    temporal patch, time step 0); any other length is kept as given.
    """
    n_space = len(data_shape) - 1
    dims = (spec,) * n_space if isinstance(spec, (int, np.integer)) else spec
    if isinstance(dims, list):
        dims = tuple(dims)
    return (*dims, data_shape[-1]) if len(dims) == n_space else dims


def _progress(progbar, n):
    """tqdm handling of base.py:230-233: None -> new bar, False -> none, else reset."""
    if progbar is False:
        return
    if progbar is None:
        try:
            from tqdm.auto import tqdm
        except ImportError:  # tqdm is optional here
            return
        progbar = tqdm(total=n)
    else:
        progbar.reset(total=n)
    progbar.update(n)
    if hasattr(progbar, "close"):
        progbar.close()


def prior_is_float32(noise_std):
    """True when the reference's prior variance array is float32 (a float32
    noise-level array squared: lowrank.py:94-97), so that it averages each
    patch's prior in float32; a float noise level gives float64."""
    if isinstance(noise_std, (float, np.floating)):
        return False
    return np.asarray(noise_std).dtype == np.float32


def spatial_variance_map(noise_std, data_shape):
    """Prior variance per voxel as the reference builds it (lowrank.py:94-101).

    A float gives a constant map; anything else is squared and broadcast over
    time (same errors as the reference for unsupported types).  Returns the
    float64 spatial map.
    """
    if isinstance(noise_std, (float, np.floating)):
        var = noise_std**2 * np.ones(data_shape[:-1])
    else:
        var = noise_std**2
    full = np.broadcast_to(var[..., None], data_shape)
    return np.ascontiguousarray(full[..., 0], dtype=np.float64)
