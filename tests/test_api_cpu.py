"""CPU tests: C-ABI library loads and exports every declared symbol, and the
host-side logic (parameters, validation, exceptions, dtype rules) mirrors the
reference.  No GPU compute is called here."""

import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "patch_denoise_amd.h")


def test_library_exports_every_declared_symbol():
    from patch_denoise import _native as N

    lib = N.load_library()
    decl = re.findall(r"^\s*(?:int|const char\*)\s+(pdamd_\w+)\s*\(", open(HEADER).read(), re.M)
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(N.EXPORTS)
    assert b"gfx950" in lib.pdamd_version()


def test_problem_struct_matches_header_fields():
    from patch_denoise import _native as N

    text = open(HEADER).read()
    body = text[text.index("typedef struct pdamd_problem {"):text.index("} pdamd_problem;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b(\w+)(?:\[4\])?;", body)
    assert [f[0] for f in N.Problem._fields_] == names


def test_workspace_size_and_validation_without_gpu():
    from patch_denoise import _native as N

    prob = N.make_problem((80, 80, 48, 300), (5, 5, 5, 300), (1, 1, 1, 300), np.float32, N.MPPCA)
    assert N.workspace_bytes(prob) > 0
    bad = N.make_problem((8, 8, 8, 20), (9, 3, 3, 20), (1, 1, 1, 20), np.float32, N.MPPCA)
    with pytest.raises(ValueError):
        N.workspace_bytes(bad)
    large = N.make_problem((20, 20, 20, 300), (6, 6, 6, 300), (1, 1, 1, 300), np.float32, N.MPPCA)
    assert N.workspace_bytes(large) > 0  # n = min(216, 300): large-Gram kernels
    big = N.make_problem((20, 20, 20, 600), (9, 9, 9, 600), (1, 1, 1, 600), np.float32, N.MPPCA)
    assert N.workspace_bytes(big) > 0  # n = min(729, 600) = 600: above 512
    huge = N.make_problem((20, 20, 20, 2500), (13, 13, 13, 2500), (1, 1, 1, 2500), np.float32,
                          N.MPPCA)
    with pytest.raises(N.NativeError, match="supported maximum"):
        N.workspace_bytes(huge)  # n = min(2197, 2500) > 2048
    short = N.make_problem((8, 8, 8, 8), (3, 3, 3, 8), (1, 1, 1, 8), np.float32, N.MPPCA)
    with pytest.raises(ValueError):
        N.workspace_bytes(short)  # eigen rules need >= 10 frames


def test_patch_params_mirror_reference():
    from oracle import cpu_ref as O
    from patch_denoise import MPPCADenoiser

    for shape, ps, po in (((10, 10, 8, 24), 3, 2), ((16, 14, 20), 6, 5),
                          ((12, 10, 6, 40), (3, 3, 3), (1, 2, 1)), ((9, 9, 7, 22), [3, 3, 3], [2, 2, 2])):
        d = MPPCADenoiser(ps, po, threshold_scale=1.0)
        assert d._get_patch_param(shape) == O.patch_params(ps, po, shape)


def test_geometry_errors_and_padding():
    from patch_denoise.space_time.base import _geometry

    assert _geometry((16, 14, 20), (6, 6, 20), (5, 5, 20)) == ((1, 16, 14, 20), (1, 6, 6, 20),
                                                                (1, 1, 1, 0))
    with pytest.raises(ValueError):
        _geometry((8, 8, 8, 20), (3, 3, 3, 20), (4, 2, 2, 20))  # overlap > patch
    with pytest.raises(ValueError):
        _geometry((8, 8, 8, 20), (3, 3, 20), (2, 2, 20))  # rank mismatch
    with pytest.raises(ValueError):
        _geometry((4, 8, 8, 20), (6, 3, 3, 20), (5, 2, 2, 20))  # patch larger than volume


def test_recombination_loss_and_nordic_errors():
    import patch_denoise as pd

    with pytest.raises(ValueError):
        pd.MPPCADenoiser(3, 2, threshold_scale=1.0, recombination="median")
    with pytest.raises(KeyError):
        pd.OptimalSVDDenoiser(3, 2, loss="max")
    data = np.zeros((6, 6, 5, 12), np.float32)
    np.random.seed(0)
    with pytest.raises(ValueError):
        pd.nordic(data, 3, 2, noise_std="not_a_float_or_array", n_iter_threshold=1)
    with pytest.raises(ValueError):
        pd.mp_pca(data, 3, 2, mask=np.ones((6, 6), bool))  # incompatible mask
    with pytest.raises(ValueError):
        pd.optimal_thresholding(data, 2, 1)  # beta = 12 / 8 > 1 (MP median)
    with pytest.raises(TypeError):
        pd.mp_pca(data.astype(np.int32), 3, 2)  # unsupported dtype


def test_nordic_threshold_matches_reference_sequence():
    from oracle import cpu_ref as O
    from patch_denoise.space_time.lowrank import nordic_threshold

    np.random.seed(3)
    a = nordic_threshold(27 * 20, 20, 4)
    np.random.seed(3)
    b = O.nordic_threshold((27 * 20,), 20, 1.0, 4)
    assert a == pytest.approx(b, rel=1e-13)
    assert isinstance(a, np.float64)


def test_mp_median_matches_oracle():
    from oracle import cpu_ref as O
    from patch_denoise.space_time.utils import marchenko_pastur_median

    for b in (0.05, 0.4, 0.96, 1.0):
        assert marchenko_pastur_median(b) == O.mp_median(b)
    with pytest.raises(ValueError):
        marchenko_pastur_median(1.2)


def test_comparison_dtype_rules():
    from patch_denoise.space_time.lowrank import _compares_in_f64

    assert not _compares_in_f64(np.float32, 2.3 ** 2)         # Python float: weak
    assert _compares_in_f64(np.float32, np.float64(5.29))      # NumPy float64: strong
    assert not _compares_in_f64(np.complex64, 1.0)
    assert _compares_in_f64(np.float64, 1.0)


def test_prior_variance_map():
    from patch_denoise.space_time.base import spatial_variance_map

    v = spatial_variance_map(2.0, (4, 5, 6, 10))
    assert v.shape == (4, 5, 6) and v.dtype == np.float64 and np.all(v == 4.0)
    m = np.arange(120, dtype=np.float32).reshape(4, 5, 6)
    np.testing.assert_array_equal(spatial_variance_map(m, (4, 5, 6, 10)), (m**2).astype(np.float64))
    with pytest.raises(TypeError):
        spatial_variance_map(1, (4, 5, 6, 10))  # a Python int fails like the reference


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asking for more exits with status 2 and says so
    (before any GPU or torch import)."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""
