"""FP64 re-check of near-threshold rule decisions (MP-PCA / Hybrid on float32
data with an FP32 eigen stage; ``pdamd_problem.refine_rel``).

* ``refine_rel = 1`` flags every patch (every comparison of the rule then has
  its sides "within 100 %"), so the whole call is redone by the FP64 chain: the
  retained ranks, noise estimates and top-10 eigenvalues must equal those of
  the FP64 eigen stage (``precision="fp64"``, same kernels) exactly, on the
  gather path (M-side), the atomic scatter path and a T-side volume, for
  MP-PCA and Hybrid; the outputs agree to the float32 rounding of the stored
  eigenvectors (rel-L2 <= 1e-6).
* ``refine_rel = 0`` re-checks nothing.
* The default (1e-6) on the 80x80x48x300 headline volume: every re-checked
  patch carries the FP64 stage's rank, and the others the FP32 stage's.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.nan_to_num(np.asarray(a, np.float64)), np.nan_to_num(np.asarray(b, np.float64))
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _run(den, data, mask, *, refine_rel, precision="auto", mode="auto", var_map=None):
    from patch_denoise import _native as N

    prob, p_s, p_o, mask_arr = den._prepare(data, mask, 50)
    vm = den._configure(prob, data, p_s)
    if var_map is not None:
        vm = var_map
    prob.precision = N.PREC[precision]
    prob.recombine_mode = N.RECOMBINE[mode]
    prob.refine_rel = refine_rel
    run = N.DeviceRun(prob, data, mask_arr, vm, want_patch_info=True).run()
    out, w = run.result(data.shape)
    return out, w, run.patch_info()


def _volume(shape, T, seed, rank=3):
    rng = np.random.default_rng(seed)
    sig = 300 + 20 * np.tensordot(rng.standard_normal((rank,) + shape),
                                  rng.standard_normal((rank, T)), axes=([0], [0]))
    return (sig + 10 * rng.standard_normal(shape + (T,))).astype(np.float32)


CASES = [
    # (name, spatial shape, frames, recombination mode): M-side gather,
    # M-side scatter, T-side (scatter)
    ("mside_gather", (14, 14, 12), 150, "gather"),
    ("mside_scatter", (14, 14, 12), 150, "scatter"),
    ("tside", (14, 14, 12), 60, "auto"),
]


@pytest.mark.parametrize("method", ["mppca", "hybrid"])
@pytest.mark.parametrize("name,shape,T,mode", CASES, ids=[c[0] for c in CASES])
def test_refine_everything_equals_fp64_stage(name, shape, T, mode, method):
    from patch_denoise import HybridPCADenoiser, MPPCADenoiser
    from patch_denoise.space_time.base import spatial_variance_map

    data = _volume(shape, T, seed=11)
    var_map = None
    if method == "mppca":
        den = MPPCADenoiser(5, 4, threshold_scale=1.0)
    else:
        den = HybridPCADenoiser(5, 4)
        sigma = 10.0 * (1 + 0.3 * np.random.default_rng(2).random(shape))
        var_map = spatial_variance_map(sigma, data.shape)
        den._var_map = var_map
    print(f"[refine] {name} {method} {shape}x{T}: refine all, FP64 stage", flush=True)
    o_all, w_all, i_all = _run(den, data, None, refine_rel=1.0, mode=mode, var_map=var_map)
    o64, w64, i64 = _run(den, data, None, refine_rel=0.0, precision="fp64", var_map=var_map)
    n = len(i_all["rank"])
    assert n <= 1024, "the case must fit one re-check list"
    assert np.all(i_all["refined"] == 1)
    assert np.array_equal(i_all["index"], i64["index"])
    assert np.array_equal(i_all["rank"], i64["rank"])
    assert np.array_equal(i_all["noise"], i64["noise"], equal_nan=True)
    assert np.array_equal(i_all["eig"], i64["eig"])
    e, ew = rel_l2(o_all, o64), rel_l2(w_all, w64)
    print(f"[refine] {name} {method}: {n} patches, out rel-L2 vs FP64 stage {e:.2e}, "
          f"weights {ew:.2e}", flush=True)
    assert e <= 1e-6 and ew <= 1e-6


def test_refine_off_rechecks_nothing():
    from patch_denoise import MPPCADenoiser

    data = _volume((14, 14, 12), 150, seed=3)
    den = MPPCADenoiser(5, 4, threshold_scale=1.0)
    _, _, info = _run(den, data, None, refine_rel=0.0)
    assert not np.any(info["refined"])


def test_refine_default_headline():
    """80x80x48x300: the re-checked patches (gap to a decision below 1e-6
    relative) carry the FP64 stage's ranks; the rest the FP32 stage's."""
    from patch_denoise import MPPCADenoiser, workloads

    data, mask = workloads.config2()
    den = MPPCADenoiser(5, 4, threshold_scale=1.0)
    print("[refine] 80x80x48x300: default re-check, FP32 only, FP64 stage", flush=True)
    _, _, ir = _run(den, data, mask, refine_rel=1e-6)
    _, _, i32 = _run(den, data, mask, refine_rel=0.0)
    _, _, i64 = _run(den, data, mask, refine_rel=0.0, precision="fp64")
    ref = ir["refined"].astype(bool)
    nref = int(ref.sum())
    d32 = int(np.sum(i32["rank"] != i64["rank"]))
    print(f"[refine] 80x80x48x300: {nref} of {len(ref)} patches re-checked; FP32 vs FP64 "
          f"stage rank differences {d32}, after re-check "
          f"{int(np.sum(ir['rank'] != i64['rank']))}", flush=True)
    assert np.array_equal(ir["rank"][ref], i64["rank"][ref])
    assert np.array_equal(ir["rank"][~ref], i32["rank"][~ref])
    assert nref < 1024
