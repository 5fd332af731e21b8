"""Python face of the compute-step functions (reference: the free functions of
Backend/cuda_includes/*.cuh called by the FourierTransform / Converts /
Rendering step classes).  Device memory and streams come from torch tensors
(plumbing only); every computation runs in libholo_hip.so.

Complex buffers are torch.complex64 (interleaved float2, the cuComplex layout).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .constants import (HV_DEPTH_8, HV_DEPTH_16, HV_DEPTH_32F, HV_IMG_ARGUMENT, HV_IMG_MODULUS,
                        HV_IMG_SQUARED_MODULUS, HV_LENS_FP64, HV_SPACE_ANGULARSP, HV_SPACE_FRESNELTR)

_DEPTH = {torch.uint8: HV_DEPTH_8, torch.uint16: HV_DEPTH_16, torch.float32: HV_DEPTH_32F}


def _s(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _p(t: torch.Tensor):
    assert t.is_cuda and t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr())


def _raw_depth(raw: torch.Tensor) -> int:
    if raw.dtype == torch.int16:  # uint16 frames carried in an int16 tensor
        return HV_DEPTH_16
    return _DEPTH[raw.dtype]


def convert(raw: torch.Tensor) -> torch.Tensor:
    """input_queue_to_input_buffer (tools_conversion.cu:120-171): (B, H, W) -> complex64."""
    out = torch.empty(raw.shape, dtype=torch.complex64, device=raw.device)
    B = raw.shape[0]
    P = raw[0].numel()
    _lib.call("hv_convert_to_complex", _p(out), _p(raw), P, B, _raw_depth(raw), _s())
    return out


def lens_angular_spectrum(nx: int, ny: int, z: float, lam: float, pixel_um: float, precision: int = HV_LENS_FP64,
                          device="cuda") -> torch.Tensor:
    """angular_spectrum_lens (angular_spectrum.cu:14-31): centred transfer function."""
    out = torch.empty((ny, nx), dtype=torch.complex64, device=device)
    step = float(torch.tensor(pixel_um, dtype=torch.float32) * torch.tensor(1e-6, dtype=torch.float32))
    _lib.call("hv_angular_spectrum_lens", _p(out), nx, ny, z, lam, step, step, precision, _s())
    return out


def lens_fresnel(width: int, height: int, lam: float, z: float, pixel_um: float, precision: int = HV_LENS_FP64,
                 device="cuda") -> torch.Tensor:
    """fresnel_transform_lens (fresnel_transform.cu:14-63)."""
    out = torch.empty((height, width), dtype=torch.complex64, device=device)
    _lib.call("hv_fresnel_transform_lens", _p(out), max(width, height), height, width, lam, z, pixel_um, precision,
              _s())
    return out


def prepare_lens(space: int, width: int, height: int, lam: float, z: float, pixel_um: float,
                 precision: int = HV_LENS_FP64, device="cuda") -> torch.Tensor:
    """The lens hv_propagate_to_ring expects (shifted H / N, or the chirp)."""
    out = torch.empty((height, width), dtype=torch.complex64, device=device)
    _lib.call("hv_prepare_lens", _p(out), space, width, height, lam, z, pixel_um, precision, _s())
    return out


def shift_corners(buf: torch.Tensor) -> None:
    """shift_corners (shift_corners.cu) in place on (..., H, W)."""
    H, W = buf.shape[-2:]
    batch = buf.numel() // (H * W)
    name = "hv_shift_corners_c64" if buf.dtype == torch.complex64 else "hv_shift_corners_f32"
    _lib.call(name, _p(buf), batch, W, H, _s())


def angular_spectrum(buf: torch.Tensor, lens_shifted: torch.Tensor) -> None:
    """angular_spectrum (angular_spectrum.cu:33-63), in place on (B, H, W)."""
    B, H, W = buf.shape
    _lib.call("hv_angular_spectrum", _p(buf), B, _p(lens_shifted), W, H, _s())


def fresnel_transform(buf: torch.Tensor, chirp: torch.Tensor) -> None:
    """fresnel_transform (fresnel_transform.cu:65-83), in place on (B, H, W)."""
    B, H, W = buf.shape
    _lib.call("hv_fresnel_transform", _p(buf), B, _p(chirp), W, H, _s())


def fft2(buf: torch.Tensor, direction: int = -1) -> None:
    """Unnormalised batched FFT2 in place on (B, H, W)."""
    B, H, W = buf.shape
    _lib.call("hv_fft2", _p(buf), B, W, H, direction, _s())


def propagate_to_ring(raw: torch.Tensor, space: int, lens, ring: torch.Tensor, slot0: int) -> None:
    """Fused conversion + space transform + transfer into ring slots."""
    B, H, W = raw.shape
    T = ring.shape[0]
    lp = _p(lens) if lens is not None else None
    _lib.call("hv_propagate_to_ring", _p(raw), _raw_depth(raw), B, W, H, space, lp, _p(ring), T, slot0, _s())


def stft(ring: torch.Tensor) -> torch.Tensor:
    """stft (stft.cu:9-15): (T, H, W) complex -> (T, H, W) spectrum."""
    out = torch.empty_like(ring)
    T = ring.shape[0]
    _lib.call("hv_stft", _p(out), _p(ring), T, ring[0].numel(), _s())
    return out


def complex_to_image(spec: torch.Tensor, img_type: int, pmin: int, pmax: int) -> torch.Tensor:
    """complex_to_modulus / squared_modulus / argument (tools_conversion.cu)."""
    out = torch.empty(spec.shape[1:], dtype=torch.float32, device=spec.device)
    name = {HV_IMG_MODULUS: "hv_complex_to_modulus", HV_IMG_SQUARED_MODULUS: "hv_complex_to_squared_modulus",
            HV_IMG_ARGUMENT: "hv_complex_to_argument"}[img_type]
    _lib.call(name, _p(out), _p(spec), pmin, pmax, spec[0].numel(), _s())
    return out


def stft_image(ring: torch.Tensor, img_type: int, pmin: int, pmax: int, minmax: bool = False):
    """Fused STFT + image type; returns (image, minmax tensor of 2 uint32 or None)."""
    out = torch.empty(ring.shape[1:], dtype=torch.float32, device=ring.device)
    mm = None
    if minmax:
        mm = torch.empty(2, dtype=torch.int32, device=ring.device)
        _lib.call("hv_minmax_init", _p(mm), _s())
    T = ring.shape[0]
    _lib.call("hv_stft_image", _p(out), _p(ring), T, ring[0].numel(), img_type, pmin, pmax,
              _p(mm) if mm is not None else None, _s())
    return out, mm


def stft_moments(ring: torch.Tensor, f_start: int, f_end: int, fps: float):
    """Moments M0, M1, M2 from the ring (STFT fused)."""
    m = [torch.empty(ring.shape[1:], dtype=torch.float32, device=ring.device) for _ in range(3)]
    T = ring.shape[0]
    _lib.call("hv_stft_moments", _p(m[0]), _p(m[1]), _p(m[2]), _p(ring), T, ring[0].numel(), f_start, f_end, fps,
              _s())
    return m


def pca_cov(ring: torch.Tensor, precision: int) -> torch.Tensor:
    """cov = H^H H (matrix_operations.cc:9-26); returns cov[i, j] (T x T)."""
    T = ring.shape[0]
    P = ring[0].numel()
    nbytes = ctypes.c_size_t()
    _lib.call("hv_pca_workspace_size", T, P, precision, ctypes.byref(nbytes))
    ws = torch.empty(max(nbytes.value, 4), dtype=torch.uint8, device=ring.device)
    cov = torch.empty((T, T), dtype=torch.complex64, device=ring.device)  # column-major buffer
    _lib.call("hv_pca_cov", _p(cov), _p(ring), T, P, precision, _p(ws), _s())
    return cov.t()  # buffer index i + j T  ->  [i, j]


def eig_hermitian(cov_ij: torch.Tensor):
    """FP64 Jacobi; cov_ij[i, j] complex64.  Returns (evals ascending, V[:, j])."""
    T = cov_ij.shape[0]
    colmajor = cov_ij.t().contiguous()  # buffer index i + j T
    evals = torch.empty(T, dtype=torch.float64, device=cov_ij.device)
    evecs = torch.empty((T, T), dtype=torch.complex128, device=cov_ij.device)
    nbytes = ctypes.c_size_t()
    _lib.call("hv_eig_workspace_size", T, ctypes.byref(nbytes))
    ws = torch.empty(max(nbytes.value, 16), dtype=torch.uint8, device=cov_ij.device)
    _lib.call("hv_eig_hermitian_f64", _p(evals), _p(evecs), _p(colmajor), T, _p(ws), _s())
    return evals, evecs.t()


def pca_image(ring: torch.Tensor, V: torch.Tensor, img_type: int, pmin: int, pmax: int) -> torch.Tensor:
    """C = H V (matrix_operations.cc:77-102) + image type + x131071 (converts.cc:50-61)."""
    T = ring.shape[0]
    colmajor = V.t().contiguous().to(torch.complex128)
    out = torch.empty(ring.shape[1:], dtype=torch.float32, device=ring.device)
    _lib.call("hv_pca_image", _p(out), _p(ring), _p(colmajor), T, ring[0].numel(), img_type, pmin, pmax, None, _s())
    return out


def pca_moments(ring: torch.Tensor, V: torch.Tensor, moment: int, f_start: int, f_end: int,
                fps: float) -> torch.Tensor:
    """C = H V (matrix_operations.cc:77-102), then moment `moment` of |C_f|^2 over
    [f_start, f_end] (pipe.cc:429-455, tools_compute.cu:38-97)."""
    T = ring.shape[0]
    colmajor = V.t().contiguous().to(torch.complex128)
    out = torch.empty(ring.shape[1:], dtype=torch.float32, device=ring.device)
    _lib.call("hv_pca_moments", _p(out), _p(ring), _p(colmajor), T, ring[0].numel(), moment, f_start, f_end, fps,
              None, _s())
    return out


def unwrap_2d(phase: torch.Tensor) -> torch.Tensor:
    """unwrap_2d (tools_unwrap.cu:269-369), raw output."""
    H, W = phase.shape
    nbytes = ctypes.c_size_t()
    _lib.call("hv_unwrap_scratch_size", W, H, ctypes.byref(nbytes))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=phase.device)
    out = torch.empty_like(phase)
    _lib.call("hv_unwrap_2d", _p(out), _p(phase), W, H, _p(ws), _s())
    return out


def rescale_unwrap(img: torch.Tensor) -> torch.Tensor:
    """rescale_float_unwrap2d (tools_conversion.cu:315-335), device min/max."""
    out = torch.empty_like(img)
    mm = torch.empty(2, dtype=torch.int32, device=img.device)
    _lib.call("hv_rescale_unwrap2d", _p(out), _p(img), img.numel(), _p(mm), _s())
    return out


def subsample(img: torch.Tensor, step: int) -> torch.Tensor:
    """subsample_frame (frame_reshape.cu:278-330)."""
    H, W = img.shape
    out = torch.empty((H // step, W // step), dtype=torch.float32, device=img.device)
    _lib.call("hv_subsample_frame_f32", _p(out), _p(img), W, H, step, _s())
    return out


def minmax(img: torch.Tensor) -> torch.Tensor:
    """Device min/max -> tensor([min, max])."""
    mm = torch.empty(2, dtype=torch.int32, device=img.device)
    out = torch.empty(2, dtype=torch.float32, device=img.device)
    _lib.call("hv_minmax_init", _p(mm), _s())
    _lib.call("hv_minmax", _p(mm), _p(img), img.numel(), _s())
    _lib.call("hv_minmax_decode", _p(out), _p(mm), _s())
    return out


def decode_minmax(mm: torch.Tensor) -> torch.Tensor:
    out = torch.empty(2, dtype=torch.float32, device=mm.device)
    _lib.call("hv_minmax_decode", _p(out), _p(mm), _s())
    return out


def percentile_range(img: torch.Tensor, lower: float, upper: float, step: int = 1, clamp: bool = True):
    """compute_percentile_xy_view + update_contrast (percentile.cu, contrast_api.cc:137-146)."""
    H, W = img.shape
    nbytes = ctypes.c_size_t()
    _lib.call("hv_percentile_scratch_size", img.numel(), ctypes.byref(nbytes))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=img.device)
    out = torch.empty(2, dtype=torch.float32, device=img.device)
    _lib.call("hv_percentile_range", _p(out), _p(img), W, H, step, lower, upper, 1 if clamp else 0, _p(ws), _s())
    return out


def percentile_range_reticle(img: torch.Tensor, lower: float, upper: float, scale: float, clamp: bool = True):
    """compute_percentile_xy_view with the contrast reticle on (percentile.cu:15-43, 150-222):
    percentiles over the centred disc of radius uint(scale * min(W, H) / 2)."""
    H, W = img.shape
    nbytes = ctypes.c_size_t()
    _lib.call("hv_percentile_scratch_size", img.numel(), ctypes.byref(nbytes))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=img.device)
    out = torch.empty(2, dtype=torch.float32, device=img.device)
    _lib.call("hv_percentile_range_reticle", _p(out), _p(img), W, H, scale, lower, upper, 1 if clamp else 0, _p(ws),
              _s())
    return out


def apply_contrast(img: torch.Tensor, mn: float, mx: float) -> None:
    """apply_contrast_correction (contrast_correction.cu:5-17), in place."""
    _lib.call("hv_apply_contrast", _p(img), img.numel(), mn, mx, _s())


def float_to_ushort(img: torch.Tensor) -> torch.Tensor:
    """float_to_ushort (tools_conversion.cu:350-387); result as int16 bits of u16."""
    out = torch.empty(img.shape, dtype=torch.int16, device=img.device)
    _lib.call("hv_float_to_ushort", _p(out), _p(img), img.numel(), _s())
    return out


def contrast_to_ushort(img: torch.Tensor, range_dev: torch.Tensor, invert: bool = False) -> torch.Tensor:
    out = torch.empty(img.shape, dtype=torch.int16, device=img.device)
    _lib.call("hv_contrast_to_ushort", _p(out), _p(img), img.numel(), _p(range_dev), 1 if invert else 0, _s())
    return out


def u16(t: torch.Tensor):
    """int16-carried u16 tensor -> numpy uint16."""
    return t.cpu().numpy().view("uint16")


def unpack(packed: torch.Tensor, pixels: int, bits: int, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """unpack_12_to_16bit / unpack_10_to_16bit (unpack.cu:34-60): `pixels`
    16-bit values (carried in an int16 tensor) from a uint8 packed stream."""
    if out is None:
        out = torch.empty(pixels, dtype=torch.int16, device=packed.device)
    fn = "hv_unpack_12_to_16bit" if bits == 12 else "hv_unpack_10_to_16bit"
    _lib.call(fn, _p(out), pixels, _p(packed), packed.numel(), _s(stream))
    return out
