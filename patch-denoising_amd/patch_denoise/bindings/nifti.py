"""Minimal NIfTI-1 single-file (.nii, .nii.gz) reader and writer.

The command line reads and writes volumes through nibabel
(reference bindings/utils.py:116-144: ``nib.load(...).get_fdata(float32)``,
``.affine``, ``Nifti1Image(data, affine).to_filename``); nibabel is not a
dependency here, so this module implements the part of the NIfTI-1 format
those calls use, from the published NIfTI-1 header layout (348-byte header,
``n+1`` magic, voxel data in Fortran order from ``vox_offset``):

* read: dims, datatype, scaling (``scl_slope`` / ``scl_inter``) and the
  voxel-to-world affine (sform when ``sform_code > 0``, else the qform
  quaternion when ``qform_code > 0``, else ``diag(pixdim)``);
* write: one file, the sform set from the affine (code 2, "aligned"),
  ``pixdim`` from the affine's column norms, no scaling.

Analyze pairs (.hdr/.img), NIfTI-2 and extensions are not read.
"""

import gzip
import struct

import numpy as np

_HDR = 348
_DTYPES = {
    2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 32: np.complex64,
    64: np.float64, 256: np.int8, 512: np.uint16, 768: np.uint32, 1024: np.int64,
    1280: np.uint64, 1792: np.complex128,
}
_CODES = {np.dtype(v): k for k, v in _DTYPES.items()}


def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def _quat_affine(b, c, d, qfac, pixdim, offset):
    a2 = 1.0 - (b * b + c * c + d * d)
    a = np.sqrt(a2) if a2 > 1e-7 else 0.0
    if a2 <= 1e-7:  # rounding: renormalise (b, c, d)
        n = np.sqrt(b * b + c * c + d * d)
        b, c, d = b / n, c / n, d / n
    R = np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b],
    ])
    z = pixdim[2] * (-1.0 if qfac < 0 else 1.0)
    A = np.eye(4)
    A[:3, :3] = R * np.array([pixdim[0], pixdim[1], z])[None, :]
    A[:3, 3] = offset
    return A


def read_header(raw):
    """Parse the header bytes; returns a dict of the fields used here."""
    if len(raw) < _HDR:
        raise ValueError("not a NIfTI-1 file: header shorter than 348 bytes")
    for e in ("<", ">"):
        if struct.unpack(e + "i", raw[:4])[0] == _HDR:
            break
    else:
        raise ValueError("not a NIfTI-1 file: sizeof_hdr is not 348")
    if raw[344:347] != b"n+1":
        raise ValueError("only single-file NIfTI-1 ('n+1' magic) is supported")
    dim = struct.unpack(e + "8h", raw[40:56])
    nd = dim[0]
    if not 1 <= nd <= 7:
        raise ValueError(f"NIfTI dim[0] = {nd} outside 1..7")
    datatype = struct.unpack(e + "h", raw[70:72])[0]
    if datatype not in _DTYPES:
        raise ValueError(f"unsupported NIfTI datatype code {datatype}")
    pixdim = struct.unpack(e + "8f", raw[76:108])
    vox_offset, slope, inter = struct.unpack(e + "3f", raw[108:120])
    qform_code, sform_code = struct.unpack(e + "2h", raw[252:256])
    qb, qc, qd, qx, qy, qz = struct.unpack(e + "6f", raw[256:280])
    srow = np.array(struct.unpack(e + "12f", raw[280:328]), np.float64).reshape(3, 4)
    if sform_code > 0:
        affine = np.vstack([srow, [0, 0, 0, 1]])
    elif qform_code > 0:
        affine = _quat_affine(qb, qc, qd, pixdim[0], pixdim[1:4], (qx, qy, qz))
    else:
        affine = np.diag([pixdim[1], pixdim[2], pixdim[3], 1.0])
    return {
        "endian": e, "shape": tuple(int(v) for v in dim[1:1 + nd]),
        "dtype": np.dtype(_DTYPES[datatype]).newbyteorder(e), "vox_offset": int(vox_offset),
        "slope": float(slope), "inter": float(inter), "affine": affine,
        "pixdim": pixdim,
    }


def load(path, dtype=np.float32):
    """Read a NIfTI-1 file: (data scaled to ``dtype`` (complex stays complex), affine)."""
    with _open(path, "rb") as f:
        raw = f.read()
    h = read_header(raw)
    n = int(np.prod(h["shape"]))
    off = max(h["vox_offset"], _HDR)
    data = np.frombuffer(raw, h["dtype"], count=n, offset=off).reshape(h["shape"], order="F")
    if np.iscomplexobj(data):
        out = data.astype(np.complex64 if np.dtype(dtype).itemsize <= 4 else np.complex128)
    else:
        out = data.astype(dtype)
    slope, inter = h["slope"], h["inter"]
    if slope != 0.0 and np.isfinite(slope):  # scl_slope 0 or NaN: no scaling
        inter = inter if np.isfinite(inter) else 0.0
        if slope != 1.0 or inter != 0.0:
            r = out.real.dtype.type
            out = out * r(slope) + r(inter)
    return np.ascontiguousarray(out), h["affine"]


def save(data, affine, path):
    """Write ``data`` (any dtype of the NIfTI-1 table) with ``affine`` to ``path``."""
    data = np.asarray(data)
    if data.dtype == np.bool_:
        data = data.astype(np.uint8)
    dt = np.dtype(data.dtype).newbyteorder("=")
    if dt not in _CODES:
        raise ValueError(f"dtype {data.dtype} has no NIfTI-1 datatype code")
    if not 1 <= data.ndim <= 7:
        raise ValueError("NIfTI-1 holds 1 to 7 dimensions")
    A = np.eye(4) if affine is None else np.asarray(affine, np.float64)
    if A.shape != (4, 4):  # the reference passes np.eye(ndim) when no affine is known
        B = np.eye(4)
        k = min(A.shape[0], 4) - 1
        B[:k, :k] = A[:k, :k]
        A = B
    hdr = bytearray(352)  # header + 4-byte empty extension flag
    struct.pack_into("<i", hdr, 0, _HDR)
    dim = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    struct.pack_into("<8h", hdr, 40, *dim)
    struct.pack_into("<2h", hdr, 70, _CODES[dt], dt.itemsize * 8)
    vox = np.sqrt((A[:3, :3] ** 2).sum(axis=0))
    pixdim = [1.0] + list(vox) + [1.0] * 4
    struct.pack_into("<8f", hdr, 76, *pixdim)
    struct.pack_into("<3f", hdr, 108, 352.0, 1.0, 0.0)
    struct.pack_into("<B", hdr, 123, 10)  # xyzt_units: mm, s
    struct.pack_into("<2h", hdr, 252, 0, 2)
    struct.pack_into("<12f", hdr, 280, *A[:3, :].ravel())
    hdr[344:348] = b"n+1\x00"
    with _open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(np.asarray(data, dt.newbyteorder("<")).tobytes(order="F"))
    return path
