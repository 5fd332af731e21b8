"""Named workloads: compute settings + synthetic source of each configuration.

Workloads (the batch sizes are the builder's choice where the configuration
does not fix them; S = B keeps every batch in the window):

  c1      256x256 u16 12-bit, B=S=T=16, angular spectrum (852 nm, 20 um,
          z=0.25 m), STFT, modulus of bins 2..7, min-max contrast to u16.
  metric  512x512 u16 12-bit, B=S=T=32, angular spectrum (852 nm, 20 um,
          z=0.25 m), STFT, modulus of bins 2..15, min-max contrast.
  c2      512x512 u8, B=S=T=32, Fresnel (785 nm, 12 um, z=0.4 m), PCA,
          squared modulus of eigenvectors 4..31, auto contrast.
  c3      1024x1024 u16 12-bit, speckle sigma 0.5 + 40 rad quadratic phase,
          angular spectrum (z=0.3 m), STFT T=64 with B=S=32, argument of bin
          2 then 2-D unwrap, percentile 0.5/99.5 % on an 8x8 subsample.
  c5      2048x2048 u16, 3 Doppler discs (f_d 4..60), angular spectrum
          (z=0.35 m), T=B=S=128, moment 0 (or 1) over bins / components
          4..60: c5_stft (STFT) and c5_pca (PCA, split-FP16 covariance).
"""
from __future__ import annotations

import copy
import ctypes

from ._lib import ComputeSettings, SynthParams, load
from .constants import (HV_CONTRAST_MINMAX, HV_CONTRAST_PERCENTILE, HV_CONTRAST_REFERENCE, HV_DEPTH_8, HV_DEPTH_16,
                        HV_IMG_ARGUMENT, HV_IMG_MODULUS, HV_IMG_MOMENTS_0, HV_IMG_MOMENTS_1, HV_IMG_SQUARED_MODULUS,
                        HV_LENS_FP64, HV_PCA_BF16_SPLIT, HV_PCA_FP16_SPLIT, HV_SPACE_ANGULARSP, HV_SPACE_FRESNELTR, HV_TIME_PCA,
                        HV_TIME_STFT)

# settings fields (hv_compute_settings) and synthetic-source fields per workload
WORKLOADS = {
    "c1": dict(
        settings=dict(width=256, height=256, depth=HV_DEPTH_16, batch_size=16, time_stride=16,
                      time_transformation_size=16, space_transformation=HV_SPACE_ANGULARSP,
                      time_transformation=HV_TIME_STFT, img_type=HV_IMG_MODULUS, lambda_=852e-9, z_distance=0.25,
                      pixel_size=20.0, p_start=2, p_width=5, contrast_mode=HV_CONTRAST_MINMAX),
        synth=dict(bits=12, sigma=0.3, mean_counts=1024.0, discs=[(60.0, 1, 7)]),
        seed=1000),
    "metric": dict(
        settings=dict(width=512, height=512, depth=HV_DEPTH_16, batch_size=32, time_stride=32,
                      time_transformation_size=32, space_transformation=HV_SPACE_ANGULARSP,
                      time_transformation=HV_TIME_STFT, img_type=HV_IMG_MODULUS, lambda_=852e-9, z_distance=0.25,
                      pixel_size=20.0, p_start=2, p_width=13, contrast_mode=HV_CONTRAST_MINMAX),
        synth=dict(bits=12, sigma=0.3, mean_counts=1024.0, discs=[(120.0, 1, 15)]),
        seed=1001),
    "c2": dict(
        settings=dict(width=512, height=512, depth=HV_DEPTH_8, batch_size=32, time_stride=32,
                      time_transformation_size=32, space_transformation=HV_SPACE_FRESNELTR,
                      time_transformation=HV_TIME_PCA, img_type=HV_IMG_SQUARED_MODULUS, lambda_=785e-9,
                      z_distance=0.4, pixel_size=12.0, p_start=4, p_width=27, contrast_mode=HV_CONTRAST_REFERENCE,
                      pca_precision=HV_PCA_BF16_SPLIT),
        synth=dict(bits=8, sigma=0.3, mean_counts=64.0, discs=[(120.0, 1, 15)]),
        seed=2000),
    "c3": dict(
        settings=dict(width=1024, height=1024, depth=HV_DEPTH_16, batch_size=32, time_stride=32,
                      time_transformation_size=64, space_transformation=HV_SPACE_ANGULARSP,
                      time_transformation=HV_TIME_STFT, img_type=HV_IMG_ARGUMENT, unwrap_2d=1, lambda_=852e-9,
                      z_distance=0.3, pixel_size=20.0, p_start=2, p_width=0, contrast_mode=HV_CONTRAST_PERCENTILE,
                      contrast_lower=0.5, contrast_upper=99.5, contrast_subsample=8),
        synth=dict(bits=12, sigma=0.5, mean_counts=1024.0, quad_phase=40.0, discs=[(240.0, 1, 7)]),
        seed=3000),
    "c5_stft": dict(
        settings=dict(width=2048, height=2048, depth=HV_DEPTH_16, batch_size=128, time_stride=128,
                      time_transformation_size=128, space_transformation=HV_SPACE_ANGULARSP,
                      time_transformation=HV_TIME_STFT, img_type=HV_IMG_MOMENTS_0, lambda_=852e-9, z_distance=0.35,
                      pixel_size=20.0, p_start=4, p_width=56, camera_fps=128.0, contrast_mode=HV_CONTRAST_MINMAX),
        synth=dict(bits=12, sigma=0.3, mean_counts=1024.0,
                   discs=[(300.0, 4, 20, 600.0, 700.0), (300.0, 21, 40, 1400.0, 800.0),
                          (300.0, 41, 60, 1000.0, 1500.0)]),
        seed=5000),
    "c5_pca": dict(
        settings=dict(width=2048, height=2048, depth=HV_DEPTH_16, batch_size=128, time_stride=128,
                      time_transformation_size=128, space_transformation=HV_SPACE_ANGULARSP,
                      time_transformation=HV_TIME_PCA, img_type=HV_IMG_MOMENTS_0, lambda_=852e-9, z_distance=0.35,
                      pixel_size=20.0, p_start=4, p_width=56, camera_fps=128.0, contrast_mode=HV_CONTRAST_MINMAX,
                      pca_precision=HV_PCA_FP16_SPLIT),
        synth=dict(bits=12, sigma=0.3, mean_counts=1024.0,
                   discs=[(300.0, 4, 20, 600.0, 700.0), (300.0, 21, 40, 1400.0, 800.0),
                          (300.0, 41, 60, 1000.0, 1500.0)]),
        seed=5000),
}


def settings_for(name: str, **overrides) -> ComputeSettings:
    """hv_compute_settings of a workload, starting from the reference defaults.
    The workloads name their steps; renormalisation (on by default, as in the
    reference) is not among them -- it scales the image by a constant that the
    contrast step removes again -- so it is off unless overridden."""
    s = ComputeSettings()
    load().hv_default_settings(ctypes.byref(s))
    for k, v in WORKLOADS[name]["settings"].items():
        setattr(s, k, v)
    s.lens_precision = HV_LENS_FP64
    s.renorm_enabled = 0
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def synth_for(name: str, seed_offset: int = 0, **overrides) -> SynthParams:
    """Synthetic interferogram source of a workload (seed = config index x 1000 + source index)."""
    w = WORKLOADS[name]
    st = w["settings"]
    sy = copy.deepcopy(w["synth"])
    p = SynthParams()
    p.width, p.height, p.depth = st["width"], st["height"], st["depth"]
    p.bits = sy["bits"]
    p.period = st["time_transformation_size"]
    p.sigma = sy["sigma"]
    p.tilt_x, p.tilt_y = 0.15, 0.1
    p.quad_phase = sy.get("quad_phase", 0.0)
    p.mean_counts = sy["mean_counts"]
    discs = sy["discs"]
    p.n_discs = len(discs)
    for i, d in enumerate(discs):
        r, lo, hi = d[0], d[1], d[2]
        cx = d[3] if len(d) > 3 else (st["width"] - 1) / 2.0
        cy = d[4] if len(d) > 4 else (st["height"] - 1) / 2.0
        p.disc_cx[i], p.disc_cy[i], p.disc_r[i] = cx, cy, r
        p.fd_lo[i], p.fd_hi[i] = lo, hi
    p.seed = w["seed"] + seed_offset
    for k, v in overrides.items():
        setattr(p, k, v)
    return p



# Serving mix: (workload, settings overrides, batches per second).  All
# sources deliver uint16 frames from the config-1 style generator.
SERVING_MIX = [
    ("c1", {}, 1500.0),                        # 256^2, B = S = T = 16
    ("metric", {}, 600.0),                     # 512^2, B = S = T = 32
    ("c2", {"depth": HV_DEPTH_16}, 300.0),     # 512^2 PCA, T = 32
    ("c3", {}, 80.0),                          # 1024^2, T = 64, B = S = 32
]
SERVING_BUDGET_BYTES = 64 * 10**9
