"""ctypes binding to ``libpdamd.so`` (C ABI in ``include/patch_denoise_amd.h``).

PyTorch is used only as device-memory and stream plumbing: buffers are torch
tensors on ``cuda:0`` whose raw pointers are handed to the C ABI, and work is
enqueued on torch's current stream.  There is no CPU fallback: if the library
or a GPU is missing, every entry point raises ``RuntimeError``.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libpdamd.so")

F32, F64, C64, C128 = 0, 1, 2, 3
MPPCA, HYBRID, HARD, OPTIMAL, ADAPTIVE = 0, 1, 2, 3, 4
ADAPT = {"sure": 0, "gsure": 1, "qut": 2}
WEIGHTED, AVERAGE, CENTER = 0, 1, 2
LOSS = {"fro": 0, "nuc": 1, "ope": 2}
RECOMB = {"weighted": WEIGHTED, "average": AVERAGE, "center": CENTER}
MASK_NONE, MASK_SPATIAL, MASK_FULL = 0, 1, 2
PREC = {"auto": 0, "fp32": 1, "fp64": 2, "bf16x3": 3}
RECOMBINE = {"auto": 0, "scatter": 1, "gather": 2}

_DTYPES = {
    np.dtype(np.float32): F32,
    np.dtype(np.float64): F64,
    np.dtype(np.complex64): C64,
    np.dtype(np.complex128): C128,
}


class Problem(ctypes.Structure):
    """Mirror of ``pdamd_problem``."""

    _fields_ = [
        ("dtype", ctypes.c_int),
        ("shape", ctypes.c_int64 * 4),
        ("patch", ctypes.c_int64 * 4),
        ("step", ctypes.c_int64 * 4),
        ("method", ctypes.c_int),
        ("recombination", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("threshold_scale2", ctypes.c_double),
        ("scale_f64", ctypes.c_int),
        ("tau", ctypes.c_double),
        ("tau_f64", ctypes.c_int),
        ("loss", ctypes.c_int),
        ("mp_median", ctypes.c_double),
        ("var_map_dev", ctypes.c_void_p),
        ("mask_dev", ctypes.c_void_p),
        ("mask_kind", ctypes.c_int),
        ("mask_threshold", ctypes.c_double),
        ("input_dev", ctypes.c_void_p),
        ("output_dev", ctypes.c_void_p),
        ("weights_dev", ctypes.c_void_p),
        ("patch_index_dev", ctypes.c_void_p),
        ("rank_dev", ctypes.c_void_p),
        ("noise_dev", ctypes.c_void_p),
        ("eig_dev", ctypes.c_void_p),
        ("selected", ctypes.c_int64),
        ("lam_dev", ctypes.c_void_p),
        ("stage_ms", ctypes.c_void_p),
        ("stage_launches", ctypes.c_void_p),
        ("adaptive_method", ctypes.c_int),
        ("n_gamma", ctypes.c_int),
        ("gamma_dev", ctypes.c_void_p),
        ("tau0_log", ctypes.c_double),
        ("qut_tau_dev", ctypes.c_void_p),
        ("launch_per_cu", ctypes.c_int * 4),
        ("recombine_mode", ctypes.c_int),
        ("refine_rel", ctypes.c_double),
        ("refined_dev", ctypes.c_void_p),
        ("var_f32", ctypes.c_int),
    ]


# Relative gap below which an MP-PCA / Hybrid decision taken on FP32
# eigenvalues is re-checked with an FP64 eigen stage (pdamd_problem.refine_rel).
REFINE_REL = 1e-6

STAGES = ("select", "gram", "trieig", "synth", "finalize", "tridiag", "adj_cross", "adj_spectral",
          "adj_synth", "recombine", "refine")


EXPORTS = (
    "pdamd_workspace_size",
    "pdamd_denoise",
    "pdamd_select_patches",
    "pdamd_batch_workspace_size",
    "pdamd_denoise_batch",
    "pdamd_backward_workspace_size",
    "pdamd_optimal_backward",
    "pdamd_noise_workspace_size",
    "pdamd_top_sv_workspace_size",
    "pdamd_top_singular_values",
    "pdamd_estimate_noise",
    "pdamd_noalloc_scratch_bytes",
    "pdamd_denoise_noalloc",
    "pdamd_last_error",
    "pdamd_version",
)

_lib = None


def load_library():
    """Load the shared library (no GPU needed just to load it)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch's own HIP runtime libraries must be in the process before this
    # library pulls in the system ones: loaded the other way round, this
    # library's HIP calls later report that no device exists.  Importing
    # torch loads them without touching the GPU.
    import torch  # noqa: F401

    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"native library {LIB_PATH} is not built; run "
            "`python patch-denoising_amd/build.py` (or __graft_entry__.build())"
        )
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(Problem)
    lib.pdamd_workspace_size.argtypes = [P, ctypes.POINTER(ctypes.c_size_t)]
    lib.pdamd_workspace_size.restype = ctypes.c_int
    lib.pdamd_denoise.argtypes = [P, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.pdamd_denoise.restype = ctypes.c_int
    lib.pdamd_select_patches.argtypes = [
        P, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_size_t,
        ctypes.c_void_p,
    ]
    lib.pdamd_select_patches.restype = ctypes.c_int
    lib.pdamd_batch_workspace_size.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
    lib.pdamd_batch_workspace_size.restype = ctypes.c_int
    lib.pdamd_denoise_batch.argtypes = [P, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_int, ctypes.c_void_p]
    lib.pdamd_denoise_batch.restype = ctypes.c_int
    lib.pdamd_backward_workspace_size.argtypes = [P, ctypes.POINTER(ctypes.c_size_t)]
    lib.pdamd_backward_workspace_size.restype = ctypes.c_int
    lib.pdamd_optimal_backward.argtypes = [
        P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_size_t, ctypes.c_void_p,
    ]
    lib.pdamd_optimal_backward.restype = ctypes.c_int
    I64P = ctypes.POINTER(ctypes.c_int64)
    lib.pdamd_noise_workspace_size.argtypes = [I64P, ctypes.POINTER(ctypes.c_size_t)]
    lib.pdamd_noise_workspace_size.restype = ctypes.c_int
    lib.pdamd_estimate_noise.argtypes = [ctypes.c_void_p, ctypes.c_int, I64P, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p]
    lib.pdamd_estimate_noise.restype = ctypes.c_int
    lib.pdamd_top_sv_workspace_size.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_size_t)]
    lib.pdamd_top_sv_workspace_size.restype = ctypes.c_int
    lib.pdamd_top_singular_values.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                              ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_size_t, ctypes.c_void_p]
    lib.pdamd_top_singular_values.restype = ctypes.c_int
    lib.pdamd_noalloc_scratch_bytes.argtypes = [ctypes.POINTER(ctypes.c_size_t)]
    lib.pdamd_noalloc_scratch_bytes.restype = ctypes.c_int
    lib.pdamd_denoise_noalloc.argtypes = [P, ctypes.c_void_p]
    lib.pdamd_denoise_noalloc.restype = ctypes.c_int
    lib.pdamd_last_error.restype = ctypes.c_char_p
    lib.pdamd_version.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("patch_denoise needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class NativeError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        msg = load_library().pdamd_last_error().decode()
        if rc in (-1,):
            raise ValueError(msg)
        raise NativeError(f"pdamd error {rc}: {msg}")


def dtype_code(dtype):
    try:
        return _DTYPES[np.dtype(dtype)]
    except KeyError:
        raise TypeError(
            f"unsupported input dtype {np.dtype(dtype)}; use float32, float64, "
            "complex64 or complex128"
        ) from None


_TORCH_DT = {F32: "float32", F64: "float64", C64: "complex64", C128: "complex128"}


def make_problem(shape4, patch4, step4, dtype, method, recombination="weighted", precision="auto"):
    p = Problem()
    p.dtype = dtype_code(dtype)
    for i in range(4):
        p.shape[i] = int(shape4[i])
        p.patch[i] = int(patch4[i])
        p.step[i] = int(step4[i])
    p.method = method
    p.recombination = RECOMB[recombination]
    p.precision = PREC[precision]
    p.mask_kind = MASK_NONE
    p.mask_threshold = 50.0
    p.tau0_log = float("nan")
    p.refine_rel = REFINE_REL
    return p


def workspace_bytes(prob):
    n = ctypes.c_size_t(0)
    _check(load_library().pdamd_workspace_size(ctypes.byref(prob), ctypes.byref(n)))
    return int(n.value)


class DeviceRun:
    """Device buffers + one ``pdamd_denoise`` call for a 4-D volume.

    ``data`` may be a NumPy array (copied to the GPU) or a torch CUDA tensor
    already resident in device memory.  Outputs stay on the device until
    :meth:`result` copies them back.
    """

    def __init__(self, prob, data, mask=None, var_map=None, want_patch_info=False,
                 workspace=True):
        torch = _torch()
        self.torch = torch
        self.prob = prob
        dev = torch.device("cuda", torch.cuda.current_device())
        tdt = getattr(torch, _TORCH_DT[prob.dtype])
        shape4 = tuple(int(v) for v in prob.shape)
        if isinstance(data, np.ndarray):
            self.x = torch.from_numpy(np.ascontiguousarray(data).reshape(shape4)).to(dev)
        else:
            self.x = data.reshape(shape4).contiguous()
        assert self.x.dtype == tdt and self.x.is_cuda
        self.out = torch.empty(shape4, dtype=tdt, device=dev)
        self.w = torch.empty(shape4, dtype=torch.float32, device=dev)
        self.mask = None
        if mask is not None:
            m = np.ascontiguousarray(mask).astype(np.uint8)
            self.mask = torch.from_numpy(m.reshape(-1)).to(dev)
            prob.mask_dev = self.mask.data_ptr()
        self.var = None
        if var_map is not None:
            v = np.ascontiguousarray(var_map, dtype=np.float64).reshape(-1)
            self.var = torch.from_numpy(v).to(dev)
            prob.var_map_dev = self.var.data_ptr()
        prob.input_dev = self.x.data_ptr()
        prob.output_dev = self.out.data_ptr()
        prob.weights_dev = self.w.data_ptr()
        self.npatch = 1
        for i in range(4):
            S, p, st = shape4[i], int(prob.patch[i]), int(prob.step[i])
            self.npatch *= ((S - p) // st + 1) if (p < S and st > 0) else 1
        self.info = None
        if want_patch_info:
            cap = max(self.npatch, 1)
            self.info = dict(
                index=torch.empty(cap, dtype=torch.int32, device=dev),
                rank=torch.empty(cap, dtype=torch.int32, device=dev),
                noise=torch.empty(cap, dtype=torch.float32, device=dev),
                eig=torch.empty(cap * 10, dtype=torch.float32, device=dev),
                refined=torch.empty(cap, dtype=torch.int32, device=dev),
            )
            prob.patch_index_dev = self.info["index"].data_ptr()
            prob.rank_dev = self.info["rank"].data_ptr()
            prob.noise_dev = self.info["noise"].data_ptr()
            prob.eig_dev = self.info["eig"].data_ptr()
            prob.refined_dev = self.info["refined"].data_ptr()
        self.ws_bytes, self.ws = 0, None
        if workspace:
            self.ws_bytes = workspace_bytes(prob)
            self.ws = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=dev)

    def run(self):
        stream = self.torch.cuda.current_stream().cuda_stream
        _check(load_library().pdamd_denoise(ctypes.byref(self.prob), self.ws.data_ptr(),
                                            self.ws_bytes, stream))
        return self

    @property
    def selected(self):
        return int(self.prob.selected)

    def result(self, out_shape):
        out = self.out.cpu().numpy().reshape(out_shape)
        w = self.w.cpu().numpy().reshape(out_shape)
        return out, w

    def patch_info(self):
        n = self.selected
        i = self.info
        return dict(
            index=i["index"][:n].cpu().numpy(),
            rank=i["rank"][:n].cpu().numpy(),
            noise=i["noise"][:n].cpu().numpy(),
            eig=i["eig"][: n * 10].cpu().numpy().reshape(n, 10),
            refined=i["refined"][:n].cpu().numpy(),
        )


def select_patches(prob, mask=None):
    """Run only the selection kernels; returns the increasing patch indices."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device())
    keep = []
    if mask is not None:
        m = torch.from_numpy(np.ascontiguousarray(mask).astype(np.uint8).reshape(-1)).to(dev)
        keep.append(m)
        prob.mask_dev = m.data_ptr()
    nb = workspace_bytes(prob)
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
    npatch = 1
    for i in range(4):
        S, p, st = int(prob.shape[i]), int(prob.patch[i]), int(prob.step[i])
        npatch *= ((S - p) // st + 1) if (p < S and st > 0) else 1
    idx = torch.empty(max(npatch, 1), dtype=torch.int32, device=dev)
    cnt = ctypes.c_int64(0)
    _check(load_library().pdamd_select_patches(
        ctypes.byref(prob), idx.data_ptr(), ctypes.byref(cnt), ws.data_ptr(), nb,
        torch.cuda.current_stream().cuda_stream))
    return idx[: cnt.value].cpu().numpy()


class BatchRun:
    """Several volumes in one ``pdamd_denoise_batch`` call.

    ``items``: list of ``(prob, data, mask, var_map)`` as for :class:`DeviceRun`
    (each problem already configured).  Device buffers are allocated once and
    reused by every :meth:`run`, so repeated runs can replay the captured graph.
    """

    def __init__(self, items, use_graph=True, want_patch_info=False):
        torch = _torch()
        self.torch = torch
        self.runs = [DeviceRun(prob, data, mask, var_map, want_patch_info=want_patch_info,
                               workspace=False)
                     for prob, data, mask, var_map in items]
        n = len(self.runs)
        self.probs = (Problem * n)()
        for i, r in enumerate(self.runs):
            ctypes.memmove(ctypes.byref(self.probs[i]), ctypes.byref(r.prob), ctypes.sizeof(Problem))
        nb = ctypes.c_size_t(0)
        _check(load_library().pdamd_batch_workspace_size(self.probs, n, ctypes.byref(nb)))
        self.ws_bytes = int(nb.value)
        dev = torch.device("cuda", torch.cuda.current_device())
        self.ws = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=dev)
        self.use_graph = bool(use_graph)

    def run(self):
        stream = self.torch.cuda.current_stream().cuda_stream
        _check(load_library().pdamd_denoise_batch(self.probs, len(self.runs), self.ws.data_ptr(),
                                                  self.ws_bytes, int(self.use_graph), stream))
        return self

    @property
    def selected(self):
        return [int(p.selected) for p in self.probs]

    def results(self, shapes):
        return [r.result(s) for r, s in zip(self.runs, shapes)]

    def patch_info(self):
        """Per-volume selected patch indices, ranks, noise and eigenvalues
        (only when constructed with ``want_patch_info``)."""
        out = []
        for r, p in zip(self.runs, self.probs):
            r.prob.selected = p.selected
            out.append(r.patch_info())
        return out


def estimate_noise(data, block_size):
    """Sliding-window noise map on the GPU (``pdamd_estimate_noise``).

    ``data``: NumPy array or CUDA tensor with 1-3 spatial axes then time.
    Returns a float64 CUDA tensor of the spatial shape, NaN where no window
    is centred.
    """
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device())
    shape = tuple(int(v) for v in data.shape)
    nsp = len(shape) - 1
    if not 1 <= nsp <= 3:
        raise ValueError("estimate_noise takes 1 to 3 spatial axes plus time")
    code = dtype_code(data.dtype if isinstance(data, np.ndarray) else
                      np.dtype(str(data.dtype).replace("torch.", "")))
    shape4 = (1,) * (3 - nsp) + shape
    if isinstance(data, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(data)).to(dev)
    else:
        x = data.contiguous()
    out = torch.full(shape4[:3], float("nan"), dtype=torch.float64, device=dev)
    s4 = (ctypes.c_int64 * 4)(*shape4)
    nb = ctypes.c_size_t(0)
    lib = load_library()
    _check(lib.pdamd_noise_workspace_size(s4, ctypes.byref(nb)))
    ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=dev)
    _check(lib.pdamd_estimate_noise(x.data_ptr(), code, s4, nsp, int(block_size), out.data_ptr(),
                                    ws.data_ptr(), int(nb.value),
                                    torch.cuda.current_stream().cuda_stream))
    return out.reshape(shape[:-1])


def top_singular_values(a):
    """Largest singular value of a float64 matrix, or of each matrix of a
    stack (``a``: (rows, cols) or (batch, rows, cols), NumPy or CUDA tensor),
    on the GPU (``pdamd_top_singular_values``).  Returns a NumPy float64 array
    of shape (batch,) (or a 0-d array for one matrix)."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device())
    if isinstance(a, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    else:
        x = a.to(device=dev, dtype=torch.float64).contiguous()
    one = x.dim() == 2
    if one:
        x = x.unsqueeze(0)
    if x.dim() != 3:
        raise ValueError("top_singular_values takes a matrix or a stack of matrices")
    batch, rows, cols = (int(v) for v in x.shape)
    lib = load_library()
    nb = ctypes.c_size_t(0)
    _check(lib.pdamd_top_sv_workspace_size(batch, rows, cols, ctypes.byref(nb)))
    ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=dev)
    sv = torch.empty(batch, dtype=torch.float64, device=dev)
    _check(lib.pdamd_top_singular_values(x.data_ptr(), batch, rows, cols, sv.data_ptr(),
                                         ws.data_ptr(), int(nb.value),
                                         torch.cuda.current_stream().cuda_stream))
    out = sv.cpu().numpy()
    return out[0] if one else out
