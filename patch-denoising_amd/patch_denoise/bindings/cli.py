"""``patch-denoise`` command line (reference bindings/cli.py:115-507).

Same arguments and order of work as the reference's ``main``: load the input
(magnitude, or magnitude + phase -> complex), optional NaN replacement,
mask (file, or computed from the mean EPI image), optional noise map, the
``--conf`` string overriding the separate parameters, one call of the
functional API, then the output and noise map written with the input's
affine.  Every method runs on the GPU (``--gpu`` is accepted and changes
nothing); the reference's HTML mask report (nilearn) is not written, the mask
itself is, as ``mask_<output name>``.

Run as ``python -m patch_denoise.bindings.cli INPUT [OUTPUT] --method ...``.
"""

import argparse
import logging
import re
import sys
from functools import partial
from pathlib import Path

import numpy as np

from .masking import compute_epi_mask
from .utils import (
    DENOISER_MAP,
    NOISE_MAP_METHODS,
    DenoiseParameters,
    load_as_array,
    load_complex_nifti,
    save_array,
)

__version__ = "0.1.0+mi355x"
DENOISER_NAMES = ", ".join(d for d in DENOISER_MAP if d)


def _is_file(path, parser):
    p = Path(path)
    if not p.exists():
        parser.error(f"Path does not exist: <{path}>.")
    if not p.is_file():
        parser.error(f"Path should point to a file (or symlink of file): <{p.absolute()}>.")
    return p.absolute()


def _is_file_or_value(path, parser, *values):
    return path if path in values else _is_file(path, parser)


def _positive_int(string, is_parser=True):
    error = argparse.ArgumentTypeError if is_parser else ValueError
    try:
        v = int(string)
    except ValueError:
        raise error("Argument must be a nonnegative integer.") from None
    if v < 0:
        raise error("Int argument must be nonnegative.")
    return v


def _tuple_positive_int(string, is_parser=True):
    """``"11"`` -> 11; ``"5x5x3"`` (any non-digit separator) -> (5, 5, 3)."""
    sep = re.search(r"\D", string)
    if sep is None:
        return _positive_int(string, is_parser)
    return tuple(_positive_int(s, is_parser) for s in string.split(sep.group(0)))


def _key_value(spec):
    """``"key=value"`` -> (key, value); the value becomes a float when it parses
    as one.  Anything but exactly one ``=`` is rejected (cli.py:91-112)."""
    key, sep, value = spec.partition("=")
    if not sep or "=" in value:
        raise ValueError("Extra arguments must be in the form key=value.")
    try:
        return key, float(value)
    except ValueError:
        return key, value


class ToDict(argparse.Action):
    """argparse action: a list of ``key=value`` strings -> dict (later keys win)."""

    def __call__(self, parser, namespace, values, option_string=None):
        setattr(namespace, self.dest, dict(_key_value(spec) for spec in values))


def get_parser():
    parser = argparse.ArgumentParser(prog="patch-denoise",
                                     formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    is_file = partial(_is_file, parser=parser)
    parser.add_argument("input_file", help="Input (noisy) file.", type=is_file)
    parser.add_argument("output_file", nargs="?", default=None, type=Path,
                        help="Output (denoised) file.\nDefault is D<input_file_name>.")
    parser.add_argument("--version", action="version", version=__version__)
    parser.add_argument("--gpu", action="store_true",
                        help="Accepted for compatibility: every method runs on the GPU.")
    g = parser.add_argument_group("Denoising parameters")
    conf = g.add_mutually_exclusive_group(required=True)
    conf.add_argument("--method", choices=DENOISER_MAP, default="optimal-fro",
                      help=f"Denoising method: {DENOISER_NAMES}. Exclusive with --conf.")
    tpi = partial(_tuple_positive_int, is_parser=True)
    g.add_argument("--patch-shape", default=11, type=tpi, metavar="INT",
                   help="Patch shape: one int, or NxMx... per dimension.")
    g.add_argument("--patch-overlap", default=5, type=tpi, metavar="INT",
                   help="Patch overlap: one int, or NxMx... per dimension.")
    g.add_argument("--recombination", default="weighted", choices=["mean", "weighted"],
                   help="Recombination of the overlapping patches.")
    g.add_argument("--mask-threshold", default=10, type=int, metavar="INT",
                   help="Minimum percentage of masked voxels for a patch to be processed.")
    conf.add_argument("--conf", default=None,
                      help="<name>_<patch-size>_<patch-overlap>_<recombination>_<mask_threshold>.")
    g.add_argument("--extra", metavar="key=value", default=None, nargs="+", action=ToDict,
                   help="extra key=value arguments for denoising methods.")
    d = parser.add_argument_group("Additional input data")
    d.add_argument("--mask", metavar="FILE|auto", default=None,
                   type=lambda p: _is_file_or_value(p, parser, "auto"),
                   help="mask file; if auto (or absent) it is computed from the mean image.")
    d.add_argument("--noise-map", metavar="FILE", default=None, type=is_file,
                   help="noise map estimation file")
    d.add_argument("--input-phase", metavar="FILE", default=None, type=is_file,
                   help="Phase of the input data, in radians.")
    d.add_argument("--noise-map-phase", metavar="FILE", default=None, type=is_file,
                   help="Phase component of the noise map, in radians.")
    m = parser.add_argument_group("Miscellaneous options")
    m.add_argument("--output-noise-map", metavar="FILE", default=None, type=Path,
                   help="Output name for calculated noise map")
    m.add_argument("--nan-to-num", metavar="VALUE", default=None, type=float,
                   help="Replace NaN by the provided value.")
    m.add_argument("-v", "--verbose", action="count", default=0,
                   help="Increase verbosity level (-v, -vv).")
    return parser


def parse_args(argv=None):
    args = get_parser().parse_args(argv)
    if not args.extra:
        args.extra = {}
    levels = [logging.WARNING, logging.INFO, logging.DEBUG]
    logging.basicConfig(level=levels[min(args.verbose, len(levels) - 1)])
    return args


def _with_prefix(path, prefix):
    path = Path(path)
    name = path.name
    return path.with_name(prefix + name)


def main(argv=None):
    """Command line entry point; returns the output path."""
    args = parse_args(argv)
    if args.input_phase is not None:
        data, affine = load_complex_nifti(args.input_file, args.input_phase)
    else:
        data, affine = load_as_array(args.input_file)
    kwargs = dict(args.extra)
    if args.nan_to_num is not None:
        data = np.nan_to_num(data, nan=args.nan_to_num)
    logging.info(f"Input data shape: {data.shape}")
    n_nans = int(np.isnan(data).sum())
    if n_nans > 0:
        logging.warning(f"{n_nans}/{data.size} voxels are NaN. "
                        "You might want to use --nan-to-num=<value>")

    if args.mask is None or args.mask == "auto":
        mask, affine_mask = compute_epi_mask(data, affine), affine
    else:
        m, affine_mask = load_as_array(args.mask)
        mask = np.asarray(m).astype(bool)

    if args.noise_map is not None and args.noise_map_phase is not None:
        noise_map, affine_nm = load_complex_nifti(args.noise_map, args.noise_map_phase)
    elif args.noise_map is not None:
        noise_map, affine_nm = load_as_array(args.noise_map)
    elif args.noise_map_phase is not None:
        raise ValueError("The phase component of the noise map has been provided, "
                         "but not the magnitude.")
    else:
        noise_map, affine_nm = None, None
    if affine is not None:
        if affine_mask is not None and not np.allclose(affine, affine_mask):
            logging.warning("Affine matrix of input and mask does not match")
        if affine_nm is not None and not np.allclose(affine, affine_nm):
            logging.warning("Affine matrix of input and noise map does not match")

    if args.conf is not None:
        p = DenoiseParameters.from_str(args.conf)
        args.method, args.patch_shape, args.patch_overlap = p.method, p.patch_shape, p.patch_overlap
        args.recombination, args.mask_threshold = p.recombination, p.mask_threshold
    logging.info(f"Current Setup {args}")

    out = args.output_file
    if out is None:
        out = _with_prefix(args.input_file, "D")
    out = Path(out)
    out.parent.mkdir(exist_ok=True, parents=True)
    save_array(mask.astype(np.uint8), affine, _with_prefix(out, "mask_"))

    if args.method in NOISE_MAP_METHODS:
        if noise_map is None:
            raise RuntimeError("A noise map must be specified for this method.")
        kwargs["noise_std"] = noise_map
    func = DENOISER_MAP.get(args.method, None)
    if func is None:
        raise ValueError(f"Method {args.method} is not supported.")
    denoised, _, noise_std_map, _ = func(
        data, patch_shape=args.patch_shape, patch_overlap=args.patch_overlap, mask=mask,
        mask_threshold=args.mask_threshold, recombination=args.recombination, **kwargs)
    save_array(denoised, affine, out)
    save_array(noise_std_map, affine, args.output_noise_map)
    return out


if __name__ == "__main__":
    main(sys.argv[1:])
