"""The command line end to end on the GPU (reference tests/test_bindings.py:52-74:
``patch-denoise in.nii out.nii --mask mask.nii --conf mp-pca_... --extra
threshold_scale=2.3`` equals the functional call).

Input: the anisotropic-patch fixture written by the reference
(tests/golden/e2e_mside_aniso.npz, 12x10x6x40 float32, patches 3x3x3,
overlap 1x2x1, mask threshold 50), stored as NIfTI-1 with an all-ones mask
file.  The output file must match the reference's own MP-PCA output
(threshold_scale 2.3, weighted) to rel-L2 1e-5 and the reference test's
tolerance (rtol 1e-6, atol 1e-2), and the in-process API call to 1e-6.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_cli_matches_reference_output(tmp_path):
    import patch_denoise as pd
    from patch_denoise.bindings import cli, nifti

    z = np.load(os.path.join(GOLDEN, "e2e_mside_aniso.npz"))
    x = z["input"]
    A = np.diag([2.0, 2.0, 2.5, 1.0])
    nifti.save(x, A, tmp_path / "noisy.nii")
    nifti.save(np.ones(x.shape[:-1], np.uint8), A, tmp_path / "mask.nii")
    ps = "x".join(str(int(v)) for v in z["patch_shape"])
    po = "x".join(str(int(v)) for v in z["patch_overlap"])
    print(f"[cli] mp-pca patch {ps} overlap {po} on {x.shape}", flush=True)
    out = cli.main([str(tmp_path / "noisy.nii"), str(tmp_path / "out.nii.gz"), "--mask",
                    str(tmp_path / "mask.nii"), "--method", "mp-pca", "--patch-shape", ps,
                    "--patch-overlap", po, "--mask-threshold", str(int(z["mask_threshold"])),
                    "--extra", "threshold_scale=2.3"])
    got, aff = nifti.load(out)
    np.testing.assert_allclose(aff, A)
    want = z["mppca_ts23_weighted__out"]
    fin = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), fin)
    e = np.linalg.norm((got - want)[fin]) / np.linalg.norm(want[fin])
    print(f"[cli]   rel-L2 vs reference output {e:.3g}", flush=True)
    assert e <= 1e-5
    np.testing.assert_allclose(got[fin], want[fin], rtol=1e-6, atol=1e-2)
    api, _, _, _ = pd.mp_pca(x, tuple(z["patch_shape"]), tuple(z["patch_overlap"]),
                             mask=np.ones(x.shape[:-1], bool),
                             mask_threshold=int(z["mask_threshold"]), threshold_scale=2.3)
    assert np.linalg.norm((got - api)[fin]) / np.linalg.norm(api[fin]) <= 1e-6
