"""Command-line bindings on the CPU: NIfTI-1 I/O, parameter strings, the
parser, the EPI mask and the dispatch of ``main`` (reference
tests/test_bindings.py:46-161 is the model; the denoiser call itself is
replaced by a recorder here and runs on the GPU in tests/test_gpu_cli.py)."""

import gzip
import struct

import numpy as np
import pytest


def _header(shape, code, bitpix, endian="<", pixdim=(1, 1, 1), vox_offset=352.0, slope=0.0,
            inter=0.0, qform=None, sform=None):
    """A NIfTI-1 header written field by field from the published layout."""
    h = bytearray(352)
    struct.pack_into(endian + "i", h, 0, 348)
    struct.pack_into(endian + "8h", h, 40, len(shape), *shape, *([1] * (7 - len(shape))))
    struct.pack_into(endian + "2h", h, 70, code, bitpix)
    struct.pack_into(endian + "8f", h, 76, *((1.0,) + tuple(pixdim) + (1.0,) * (7 - len(pixdim))))
    struct.pack_into(endian + "3f", h, 108, vox_offset, slope, inter)
    if qform:
        struct.pack_into(endian + "h", h, 252, 1)
        struct.pack_into(endian + "6f", h, 256, *qform[0], *qform[1])
        struct.pack_into(endian + "f", h, 76, qform[2])  # qfac in pixdim[0]
    if sform is not None:
        struct.pack_into(endian + "h", h, 254, 1)
        struct.pack_into(endian + "12f", h, 280, *np.asarray(sform, np.float32)[:3].ravel())
    h[344:348] = b"n+1\x00"
    return bytes(h)


def _indexed(shape):
    g = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    return sum(gi * 10 ** i for i, gi in enumerate(g))  # value at [i, j, k] = i + 10 j + 100 k


@pytest.mark.parametrize("endian", ["<", ">"])
def test_nifti_reads_handmade_header(tmp_path, endian):
    from patch_denoise.bindings import nifti

    shape = (4, 3, 2)
    want = _indexed(shape).astype(np.int16)
    # quaternion (b, c, d) = (0, 0, 1): rotation by pi about z; qfac -1 flips k
    raw = _header(shape, 4, 16, endian, pixdim=(2.0, 3.0, 4.0), slope=0.5, inter=-1.0,
                  qform=((0.0, 0.0, 1.0), (10.0, 20.0, 30.0), -1.0))
    path = tmp_path / "a.nii.gz"
    with gzip.open(path, "wb") as f:
        f.write(raw + want.astype(np.dtype(np.int16).newbyteorder(endian)).tobytes(order="F"))
    got, aff = nifti.load(path)
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, want * 0.5 - 1.0)
    np.testing.assert_allclose(aff, [[-2, 0, 0, 10], [0, -3, 0, 20], [0, 0, -4, 30], [0, 0, 0, 1]])


def test_nifti_sform_wins_and_writer_layout(tmp_path):
    from patch_denoise.bindings import nifti

    shape = (3, 2, 2, 5)
    A = np.array([[0, 2, 0, 1], [1.5, 0, 0, 2], [0, 0, 3, 3], [0, 0, 0, 1.0]])
    data = _indexed(shape).astype(np.float32)
    raw = _header(shape, 16, 32, sform=A, qform=((0.0, 0.0, 1.0), (0, 0, 0), 1.0))
    (tmp_path / "s.nii").write_bytes(raw + data.tobytes(order="F"))
    got, aff = nifti.load(tmp_path / "s.nii")
    np.testing.assert_array_equal(got, data)
    np.testing.assert_allclose(aff, A)
    # the writer: datatype code, Fortran order from byte 352, sform code 2
    x = (np.arange(60).reshape(3, 4, 5) * 1.5).astype(np.float64)
    nifti.save(x, A, tmp_path / "w.nii")
    b = (tmp_path / "w.nii").read_bytes()
    assert struct.unpack("<i", b[:4])[0] == 348 and b[344:347] == b"n+1"
    assert struct.unpack("<2h", b[70:74]) == (64, 64)
    assert struct.unpack("<h", b[254:256])[0] == 2
    np.testing.assert_array_equal(np.frombuffer(b, "<f8", offset=352), x.ravel(order="F"))
    np.testing.assert_allclose(struct.unpack("<3f", b[80:92]), [1.5, 2.0, 3.0])
    with pytest.raises(ValueError):
        nifti.read_header(b[:100])


def test_load_save_array_formats(tmp_path):
    from patch_denoise.bindings import load_as_array, save_array

    x = np.random.default_rng(0).standard_normal((5, 4, 3, 6)).astype(np.float32)
    for name in ("o.npy", "o.nii", "o.nii.gz"):
        save_array(x, None, tmp_path / name)
        y, aff = load_as_array(tmp_path / name)
        np.testing.assert_array_equal(y, x)
        assert (aff is None) == name.endswith(".npy")
    with pytest.raises(ValueError):
        load_as_array(tmp_path / "o.txt")
    assert save_array(x, None, None) is None


def test_denoise_parameters():
    """Reference tests/test_bindings.py:77-82 and :111-124."""
    from patch_denoise.bindings import DenoiseParameters

    d = DenoiseParameters("optimal-fro", 11, 10, "weighted", 10)
    assert DenoiseParameters.from_str(str(d)) == d
    assert d.pretty_par == "11_10w"
    s = "optimal-fro_11_10_weighted_10"
    assert DenoiseParameters.from_str(s).pretty_name == s
    p = DenoiseParameters.from_str("mp-pca_6_5_weighted")
    assert (p.method, p.patch_shape, p.patch_overlap, p.recombination, p.mask_threshold) == \
        ("mp-pca", 6, 5, "weighted", 10)
    assert DenoiseParameters.from_str("noisy").pretty_name == "noisy"


def test_parser():
    from patch_denoise.bindings import cli

    p = cli.get_parser()
    with pytest.raises(SystemExit) as e:
        p.parse_args(["--help"])
    assert e.value.code == 0
    a = p.parse_args([__file__, "o.nii", "--method", "nordic", "--patch-shape", "5x5x3",
                      "--patch-overlap", "2", "--extra", "threshold_scale=2.3", "loss=nuc"])
    assert a.patch_shape == (5, 5, 3) and a.patch_overlap == 2
    assert a.extra == {"threshold_scale": 2.3, "loss": "nuc"}
    with pytest.raises(SystemExit):  # --method and --conf are exclusive
        p.parse_args([__file__, "--method", "raw", "--conf", "raw_5_2"])
    with pytest.raises(SystemExit):  # the input must exist
        p.parse_args(["/nonexistent/x.nii", "--method", "raw"])


def test_epi_mask_keeps_the_largest_bright_component():
    from patch_denoise.bindings.masking import compute_epi_mask

    rng = np.random.default_rng(1)
    g = np.meshgrid(*[np.arange(s, dtype=float) for s in (40, 36, 20)], indexing="ij")
    brain = ((g[0] - 20) / 13) ** 2 + ((g[1] - 18) / 11) ** 2 + ((g[2] - 10) / 7) ** 2 <= 1
    blob = ((g[0] - 3) ** 2 + (g[1] - 3) ** 2 + (g[2] - 3) ** 2) <= 4  # small bright speck
    img = 5 + rng.standard_normal((40, 36, 20))
    img[brain] += 200
    img[blob] += 200
    data = np.repeat(img[..., None], 4, axis=-1) + rng.standard_normal((40, 36, 20, 4))
    m = compute_epi_mask(data.astype(np.float32), np.diag([2.0, 2.0, 2.0, 1.0]))
    assert m.shape == brain.shape and m.dtype == bool
    iou = (m & brain).sum() / (m | brain).sum()
    assert iou > 0.9 and not m[blob].any()


def test_main_dispatch(tmp_path, monkeypatch):
    """``main`` routes the arguments to one functional-API call (recorded
    here instead of run) and writes output, mask and noise map."""
    from patch_denoise.bindings import cli, nifti, utils

    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 7, 6, 10)).astype(np.float32)
    A = np.diag([2.0, 2.0, 3.0, 1.0])
    nifti.save(x, A, tmp_path / "in.nii")
    mask = np.ones(x.shape[:-1], np.uint8)
    mask[0] = 0
    nifti.save(mask, A, tmp_path / "mask.nii")
    nifti.save(np.full(x.shape[:-1], 0.7, np.float32), A, tmp_path / "nm.nii")
    calls = []

    def fake(data, **kw):
        calls.append((data, kw))
        return data * 2, None, np.full(data.shape[:-1], 3.0, np.float32), None

    monkeypatch.setitem(utils.DENOISER_MAP, "mp-pca", fake)
    monkeypatch.setitem(utils.DENOISER_MAP, "nordic", fake)
    out = cli.main([str(tmp_path / "in.nii"), str(tmp_path / "out" / "d.nii"), "--mask",
                    str(tmp_path / "mask.nii"), "--conf", "mp-pca_6_5_weighted",
                    "--extra", "threshold_scale=2.3", "--output-noise-map",
                    str(tmp_path / "nmap.npy")])
    data, kw = calls[-1]
    np.testing.assert_array_equal(data, x)
    assert (kw["patch_shape"], kw["patch_overlap"], kw["recombination"],
            kw["mask_threshold"], kw["threshold_scale"]) == (6, 5, "weighted", 10, 2.3)
    np.testing.assert_array_equal(kw["mask"], mask.astype(bool))
    y, aff = nifti.load(out)
    np.testing.assert_array_equal(y, 2 * x)
    np.testing.assert_allclose(aff, A)
    assert (tmp_path / "out" / "mask_d.nii").exists()
    np.testing.assert_array_equal(np.load(tmp_path / "nmap.npy"), 3.0)
    # methods that need a noise map get it as noise_std, or fail without it
    cli.main([str(tmp_path / "in.nii"), str(tmp_path / "n.nii"), "--method", "nordic",
              "--noise-map", str(tmp_path / "nm.nii"), "--mask", str(tmp_path / "mask.nii")])
    np.testing.assert_array_equal(calls[-1][1]["noise_std"], np.float32(0.7))
    assert calls[-1][1]["patch_shape"] == 11 and calls[-1][1]["patch_overlap"] == 5
    with pytest.raises(RuntimeError, match="noise map"):
        cli.main([str(tmp_path / "in.nii"), str(tmp_path / "n.nii"), "--method", "nordic"])


def test_entrypoint_help():
    """Reference tests/test_bindings.py:46-49 (``patch-denoise --help`` exits 0)."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, PYTHONPATH=os.path.join(root, "patch-denoising_amd"))
    r = subprocess.run([sys.executable, "-m", "patch_denoise.bindings.cli", "--help"],
                       env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "--patch-overlap" in r.stdout


def test_main_complex_and_nan(tmp_path, monkeypatch):
    """Magnitude + phase files become one complex input; --nan-to-num replaces NaNs."""
    from patch_denoise.bindings import cli, nifti, utils

    rng = np.random.default_rng(3)
    mag = (5 + rng.random((6, 6, 5, 8))).astype(np.float32)
    mag[0, 0, 0, 0] = np.nan
    ph = rng.uniform(-np.pi, np.pi, mag.shape).astype(np.float32)
    for name, arr in (("m.nii", mag), ("p.nii", ph), ("nm.nii", np.ones(mag.shape[:-1], np.float32)),
                      ("mask.nii", np.ones(mag.shape[:-1], np.uint8))):
        nifti.save(arr, np.eye(4), tmp_path / name)
    calls = []

    def fake(data, **kw):
        calls.append((data, kw))
        return data, None, None, None

    monkeypatch.setitem(utils.DENOISER_MAP, "nordic", fake)
    out = cli.main([str(tmp_path / "m.nii"), str(tmp_path / "o.nii"), "--method", "nordic",
                    "--input-phase", str(tmp_path / "p.nii"), "--noise-map",
                    str(tmp_path / "nm.nii"), "--mask", str(tmp_path / "mask.nii"),
                    "--nan-to-num", "0"])
    data, kw = calls[-1]
    assert np.iscomplexobj(data) and not np.isnan(data).any()
    want = np.nan_to_num(mag * np.exp(1j * ph), nan=0.0)
    np.testing.assert_allclose(data, want, rtol=1e-6, atol=1e-6)
    got, _ = nifti.load(out)
    assert np.iscomplexobj(got)
