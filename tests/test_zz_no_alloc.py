"""pdamd_denoise_noalloc (fixed device scratch, no allocation) against
pdamd_denoise (caller workspace), through the C ABI.

Two calls in a row share the GPU's scratch array: the headline 80x80x48x300
MP-PCA volume (processed in scratch-sized chunks), then a smaller Hybrid-PCA
volume with a mask.  Each must equal the allocating entry point's result:
identical selected patches and ranks, output and weights within rel-L2 1e-6
(only the order / precision of the additions of overlapping patches differs).  A
problem whose per-volume arrays alone exceed the scratch is refused before
any GPU work.  Where the gather recombination's per-call factor buffers do
not fit the scratch, the no-allocation path recombines with the atomic
scatter: results then agree to float32 rounding of the weight sums (the
gather sums them in patch order in float32, the scatter in float64).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.nan_to_num(a.astype(np.float64)), np.nan_to_num(b.astype(np.float64))
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _problem(den, data, mask):
    prob, p_s, p_o, mask_arr = den._prepare(data, mask, 50)
    var_map = den._configure(prob, data, p_s)
    return prob, mask_arr, var_map


def _both(den, data, mask):
    import torch

    from patch_denoise import _native as N

    lib = N.load_library()
    prob, mask_arr, var_map = _problem(den, data, mask)
    ref = N.DeviceRun(prob, data, mask_arr, var_map, want_patch_info=True).run()
    ref_out, ref_w = ref.result(data.shape)
    ref_info = ref.patch_info()
    prob2, mask_arr, var_map = _problem(den, data, mask)
    run = N.DeviceRun(prob2, data, mask_arr, var_map, want_patch_info=True, workspace=False)
    stream = torch.cuda.current_stream().cuda_stream
    N._check(lib.pdamd_denoise_noalloc(ctypes.byref(prob2), stream))
    out, w = run.result(data.shape)
    return (out, w, run.patch_info()), (ref_out, ref_w, ref_info)


def _compare(tag, got, ref):
    out, w, info = got
    rout, rw, rinfo = ref
    assert np.array_equal(info["index"], rinfo["index"]), tag
    assert np.array_equal(info["rank"], rinfo["rank"]), tag
    assert np.array_equal(np.isnan(out), np.isnan(rout)), tag
    e, ew = rel_l2(out, rout), rel_l2(w, rw)
    print(f"[noalloc] {tag}: {len(info['index'])} patches, rel-L2 out {e:.2e} weights {ew:.2e}",
          flush=True)
    assert e <= 1e-6 and ew <= 1e-6, (tag, e, ew)


def test_noalloc_matches_allocating_entry_point():
    from patch_denoise import HybridPCADenoiser, MPPCADenoiser, _native as N, workloads
    from patch_denoise.space_time.base import spatial_variance_map

    nb = ctypes.c_size_t(0)
    N._check(N.load_library().pdamd_noalloc_scratch_bytes(ctypes.byref(nb)))
    assert nb.value == 256 << 20
    # 1: the headline volume, larger than one scratch-sized chunk
    data, mask = workloads.config2()
    print("[noalloc] call 1: mp_pca 80x80x48x300", flush=True)
    got, ref = _both(MPPCADenoiser(5, 4, threshold_scale=1.0), data, mask)
    _compare("mp_pca 80x80x48x300", got, ref)
    # 2: a smaller volume and another rule, right after (same scratch)
    rng = np.random.default_rng(3)
    shape, T = (24, 20, 16), 60
    sig = 300 + 20 * np.tensordot(rng.standard_normal((3,) + shape), rng.standard_normal((3, T)),
                                  axes=([0], [0]))
    small = (sig + 10 * rng.standard_normal(shape + (T,))).astype(np.float32)
    smask = workloads._ellipsoid(shape, (10, 9, 7))
    sigma = 10.0 * (1 + 0.2 * rng.random(shape))
    den = HybridPCADenoiser(3, 2)
    den._var_map = spatial_variance_map(sigma, small.shape)
    print("[noalloc] call 2: hybrid_pca 24x20x16x60", flush=True)
    got, ref = _both(den, small, smask)
    _compare("hybrid_pca 24x20x16x60", got, ref)


def test_noalloc_refuses_oversized_problem():
    """Temporal patches (a step along time) need float64 weight sums per
    element: 80x80x48x300 x 8 B > 256 MiB, refused before any GPU work."""
    from patch_denoise import _native as N

    prob = N.make_problem((80, 80, 48, 300), (5, 5, 5, 100), (1, 1, 1, 50), np.float32, N.MPPCA)
    prob.input_dev = prob.output_dev = prob.weights_dev = 1  # never dereferenced
    rc = N.load_library().pdamd_denoise_noalloc(ctypes.byref(prob), None)
    assert rc == -2
    assert b"no-allocation scratch" in N.load_library().pdamd_last_error()
