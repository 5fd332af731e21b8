"""holovibes_amd -- MI355X-native implementation of the Holovibes per-batch
compute path (conversion -> angular spectrum / Fresnel -> STFT / PCA ->
image type -> contrast -> u16), behind the C-ABI of include/holo_hip.h.

The compute runs in libholo_hip.so (hand-written HIP for gfx950, built by
holovibes_amd/build.py); this package only binds it.
"""
from . import constants  # noqa: F401
from ._lib import ComputeSettings, HoloError, SynthParams, load  # noqa: F401

__all__ = ["constants", "ComputeSettings", "SynthParams", "HoloError", "load"]
