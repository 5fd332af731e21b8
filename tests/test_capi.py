"""CPU checks of the C-ABI library: it loads without a GPU, exports every
symbol include/holo_hip.h declares, validates settings like the reference
setters, and its host-side generator is deterministic.  No compute calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    import holovibes_amd
    return holovibes_amd.load()


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "holo_hip.h")).read()
    return sorted(set(re.findall(r"^(?:hv_status|const char\*|void)\s+\*?(hv_[a-z0-9_]+)\(", text, re.M)))


def test_header_symbols_exported(lib):
    names = declared_symbols()
    assert len(names) > 50
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    from holovibes_amd._lib import SIGNATURES
    assert sorted(SIGNATURES) == declared_symbols()


def test_settings_struct_layout(lib):
    """The Python mirror of hv_compute_settings has the C layout: the library
    fills the reference defaults through it."""
    from holovibes_amd._lib import ComputeSettings
    s = ComputeSettings()
    lib.hv_default_settings(ctypes.byref(s))
    # Backend/includes/core/holovibes.hh:359-447
    assert (s.batch_size, s.time_stride, s.time_transformation_size) == (1, 1, 1)
    assert abs(s.lambda_ - 852e-9) < 1e-12 and s.pixel_size == 12.0 and s.z_distance == 0.0
    assert abs(s.contrast_lower - 0.02) < 1e-7 and abs(s.contrast_upper - 99.98) < 1e-4
    assert s.camera_fps == 0.0 and s.fft_shift == 0 and s.unwrap_2d == 0
    # the last fields, written by the library at their C offsets: RenormEnabled{true},
    # RenormConstant 5, ContrastReticleDisplayEnabled{false}, ContrastReticleScale{0.5}
    assert (s.renorm_enabled, s.renorm_constant, s.img_accu_level, s.log_enabled) == (1, 5, 1, 0)
    assert (s.contrast_reticle_enabled, s.contrast_reticle_scale) == (0, 0.5)
    assert s.output_buffer_size == 1024  # OutputBufferSize{1024} (holovibes.hh:375)
    assert ctypes.sizeof(ComputeSettings) == 41 * 4


def test_validate_settings(lib):
    from holovibes_amd import configs
    from holovibes_amd.constants import HV_IMG_MOMENTS_0, HV_TIME_NONE
    ok = configs.settings_for("metric")
    assert lib.hv_validate_settings(ctypes.byref(ok)) == 0
    for name in configs.WORKLOADS:
        assert lib.hv_validate_settings(ctypes.byref(configs.settings_for(name))) == 0, name
    # frame sides: powers of two or other even 2^a 3^b 5^c in [16, 4096]
    for good in (dict(width=1536), dict(width=1000), dict(width=1350), dict(width=1920, height=1080),
                 dict(height=768)):
        s = configs.settings_for("metric", **good)
        assert lib.hv_validate_settings(ctypes.byref(s)) == 0, good
    bad = [dict(time_stride=48), dict(time_stride=16), dict(batch_size=0), dict(width=504), dict(width=4100),
           dict(width=135),
           dict(time_transformation_size=300), dict(p_start=2, p_width=30), dict(depth=3), dict(lambda_=0.0),
           dict(contrast_mode=9), dict(contrast_subsample=3),
           dict(contrast_reticle_enabled=1, contrast_reticle_scale=1.5),
           dict(contrast_reticle_enabled=1, contrast_subsample=2),
           dict(img_type=HV_IMG_MOMENTS_0, time_transformation=HV_TIME_NONE)]
    for over in bad:
        s = configs.settings_for("metric", **over)
        assert lib.hv_validate_settings(ctypes.byref(s)) != 0, over
        assert lib.hv_last_error()
    # windows that are not powers of two: any length for the STFT, below 32 frames for PCA
    for T, want_ok in ((24, True), (100, True), (1, True)):
        s = configs.settings_for("metric", batch_size=T, time_stride=T, time_transformation_size=T, p_start=0,
                                 p_width=0)
        assert (lib.hv_validate_settings(ctypes.byref(s)) == 0) == want_ok, T
    for T, want_ok in ((24, True), (31, True), (48, False), (96, False), (64, True)):
        s = configs.settings_for("c2", batch_size=T, time_stride=T, time_transformation_size=T, p_width=T - 5)
        assert (lib.hv_validate_settings(ctypes.byref(s)) == 0) == want_ok, T


def test_step_functions_reject_bad_arguments_without_gpu(lib):
    # validation happens before any device work
    assert lib.hv_convert_to_complex(None, None, 16, 1, 2, None) == 1
    assert lib.hv_fft2(ctypes.c_void_p(16), 1, 56, 64, -1, None) == 2  # 56 = 8 x 7
    assert lib.hv_fft2(ctypes.c_void_p(16), 1, 45, 64, -1, None) == 2  # odd
    assert lib.hv_stft_image(ctypes.c_void_p(16), ctypes.c_void_p(16), 16, 256, 0, 3, 16, None, None) == 1
    assert lib.hv_stft_image(ctypes.c_void_p(16), ctypes.c_void_p(16), 300, 256, 0, 3, 5, None, None) == 2
    assert b"pmax" in lib.hv_last_error() or b"range" in lib.hv_last_error() or lib.hv_last_error()
    # the multi-batch push checks its arguments before touching the pipe
    assert lib.hv_pipe_push_batches(None, None, 1, None) == 1
    assert lib.hv_pipe_push_batches(ctypes.c_void_p(16), None, -1, None) == 1


def test_memory_accounting(lib):
    from holovibes_amd import configs
    s = configs.settings_for("metric")
    n = ctypes.c_size_t()
    assert lib.hv_pipe_memory_bytes(ctypes.byref(s), ctypes.byref(n)) == 0
    P = 512 * 512
    ring = 32 * P * 8
    outq = 1024 * P * 2  # OutputBufferSize u16 images
    assert ring + outq < n.value < ring + outq + 32 * P * 2 + 16 * P + 2 ** 20
    s.output_buffer_size = 0
    assert lib.hv_pipe_memory_bytes(ctypes.byref(s), ctypes.byref(n)) == 0
    assert ring < n.value < ring + 32 * P * 2 + 16 * P + 2 ** 20


def test_memory_accounting_raw_ring(lib):
    """A sliding window (B < T) on the bins-first path adds the raw-frame
    ring, T frames of `depth` bytes, to the pipe's memory; a window replaced
    by every batch, or a cycle without a time transform, has none."""
    from holovibes_amd import configs

    def nbytes(**over):
        s = configs.settings_for("c3", **over)
        n = ctypes.c_size_t()
        assert lib.hv_pipe_memory_bytes(ctypes.byref(s), ctypes.byref(n)) == 0
        return n.value

    P, T = 1024 * 1024, 64
    assert nbytes() - nbytes(time_transformation=0) == T * P * 2
    assert nbytes(depth=1) - nbytes(depth=1, time_transformation=0) == T * P
    # B = T: the bins are taken from the batch
    assert nbytes(batch_size=64, time_stride=64) - nbytes(batch_size=64, time_stride=64, time_transformation=0) == 0


def test_synth_deterministic_and_calibrated():
    from holovibes_amd import configs
    from holovibes_amd.pipe import synth_frames
    sp = configs.synth_for("c1")
    a = synth_frames(sp, 0, 8)
    b = np.concatenate([synth_frames(sp, 0, 3), synth_frames(sp, 3, 5)])
    assert np.array_equal(a, b)
    assert abs(a.mean() - 1024) < 10 and a.max() <= 4095
    sp8 = configs.synth_for("c2")
    c = synth_frames(sp8, 0, 2)
    assert c.dtype == np.uint8 and abs(c.mean() - 64) < 2
    other = configs.synth_for("c1", seed_offset=1)
    assert not np.array_equal(a[:1], synth_frames(other, 0, 1))


def test_synth_doppler_modulation():
    """Inside the disc the fringe pattern repeats with the configured Doppler
    frequency: the frame-to-frame spectrum peaks at integer bins 1..7."""
    from holovibes_amd import configs
    from holovibes_amd.pipe import synth_frames
    sp = configs.synth_for("c1", mean_counts=60000.0, bits=16)
    fr = synth_frames(sp, 0, 16).astype(np.float64)
    spec = np.abs(np.fft.fft(fr - fr.mean(0), axis=0))
    inside = spec[:, 120:136, 120:136].mean(axis=(1, 2))
    outside = spec[:, 0:16, 0:16].mean(axis=(1, 2))
    assert inside[1:8].mean() > 3 * outside[1:8].mean()
