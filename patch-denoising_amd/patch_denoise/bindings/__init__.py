"""Command-line bindings: ``python -m patch_denoise.bindings.cli`` (reference bindings/)."""

from .utils import DENOISER_MAP, DenoiseParameters, load_as_array, save_array

__all__ = ["DENOISER_MAP", "DenoiseParameters", "load_as_array", "save_array"]
