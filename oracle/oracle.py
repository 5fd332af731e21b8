"""Python face of the CPU oracle (oracle/holo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the CPU-baseline leg of bench.py, always as the
checker.  The product package (holovibes_amd) never imports it.

Complex arrays are numpy complex128 (FP64 oracle) or complex64 (FP32
baseline); frames are (B, H, W); the time window is (T, H, W) in ring slot
order.  Reference citations are in the C++ source next to each routine.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

HV_DEPTH_8, HV_DEPTH_16, HV_DEPTH_32F = 1, 2, 4
HV_SPACE_NONE, HV_SPACE_FRESNELTR, HV_SPACE_ANGULARSP = 0, 1, 2
HV_TIME_NONE, HV_TIME_STFT, HV_TIME_PCA = 0, 1, 2
HV_IMG_MODULUS, HV_IMG_SQUARED_MODULUS, HV_IMG_ARGUMENT = 0, 1, 2
HV_IMG_MOMENTS_0, HV_IMG_MOMENTS_1, HV_IMG_MOMENTS_2 = 5, 6, 7
HV_LENS_FP64, HV_LENS_REFERENCE_FP32 = 0, 1


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the Makefile beside this file."""
    src = os.path.join(_HERE, "holo_oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.or_pipe_create.restype = ctypes.c_void_p
        _lib.or_pipe_create.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.or_pipe_push.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_pipe_destroy.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.or_pipe_ring_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_pipe_enqueue_raw.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        _lib.or_pipe_ring_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_pipe_moments.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3
        _lib.or_pipe_pca_state.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3
        for name in ("or_lens_as",):
            getattr(_lib, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_void_p]
        _lib.or_lens_fresnel.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_int, ctypes.c_void_p]
        _lib.or_moments_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_fft_freqs.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_void_p]
        _lib.or_apply_contrast_f32.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_float, ctypes.c_float]
        for name in ("or_convert_f64", "or_convert_f32"):
            getattr(_lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                            ctypes.c_void_p]
        _lib.or_stft_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        _lib.or_image_type_f64.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p]
        _lib.or_pca_cov_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        _lib.or_pca_project_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                            ctypes.c_void_p]
        _lib.or_percentile_f32.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_void_p]
        _lib.or_float_to_u16.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        _lib.or_reticle_pixels_f32.restype = ctypes.c_uint64
        _lib.or_reticle_pixels_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                               ctypes.c_void_p]
        _lib.or_rescale_unwrap_f32.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def max_threads() -> int:
    return int(lib().or_max_threads())


# --------------------------------------------------------------------------
# step functions
# --------------------------------------------------------------------------
def convert(raw: np.ndarray) -> np.ndarray:
    """u8/u16/f32 frames (B, H, W) -> complex128 (tools_conversion.cu:107-171)."""
    raw = np.ascontiguousarray(raw)
    depth = {np.dtype(np.uint8): HV_DEPTH_8, np.dtype(np.uint16): HV_DEPTH_16,
             np.dtype(np.float32): HV_DEPTH_32F}[raw.dtype]
    out = np.empty(raw.shape, np.complex128)
    B = raw.shape[0]
    P = raw[0].size
    lib().or_convert_f64(_p(raw), depth, P, B, _p(out))
    return out


def lens_as(nx: int, ny: int, z: float, lam: float, pixel_um: float, precision: int = HV_LENS_FP64,
            shifted: bool = True) -> np.ndarray:
    """Angular-spectrum transfer function (masks.cu:31-59), corner-shifted
    like fourier_transform.cc:121-134 when `shifted`."""
    out = np.empty((ny, nx), np.complex128)
    step = np.float32(np.float32(pixel_um) * np.float32(1e-6))
    lib().or_lens_as(nx, ny, z, lam, float(step), float(step), precision, _p(out))
    if shifted:
        lib().or_shift_corners_c128(_p(out), 1, nx, ny)
    return out


def lens_fresnel(width: int, height: int, lam: float, z: float, pixel_um: float,
                 precision: int = HV_LENS_FP64) -> np.ndarray:
    """Fresnel chirp (masks.cu:7-29, fresnel_transform.cu:14-63)."""
    out = np.empty((height, width), np.complex128)
    lib().or_lens_fresnel(max(width, height), height, width, lam, z, pixel_um, precision, _p(out))
    return out


def shift_corners(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a).copy()
    if a.dtype == np.complex128:
        lib().or_shift_corners_c128(_p(a), a.size // (a.shape[-1] * a.shape[-2]), a.shape[-1], a.shape[-2])
    elif a.dtype == np.float32:
        lib().or_shift_corners_f32(_p(a), a.size // (a.shape[-1] * a.shape[-2]), a.shape[-1], a.shape[-2])
    else:
        raise TypeError(a.dtype)
    return a


def fft2(a: np.ndarray, direction: int = -1) -> np.ndarray:
    """Batched unnormalised 2-D FFT of (..., H, W) complex128 (own radix-2;
    mixed radix for other lengths)."""
    a = np.ascontiguousarray(a, np.complex128).copy()
    H, W = a.shape[-2:]
    if H * W == 0:
        raise ValueError("unsupported size")
    B = a.size // (H * W)
    rc = lib().or_fft2_f64(_p(a), B, W, H, direction)
    if rc:
        raise ValueError("unsupported size")
    return a


def fft1(a: np.ndarray, direction: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(a, np.complex128).copy()
    if lib().or_fft1_f64(_p(a), a.size, direction):
        raise ValueError("unsupported size")
    return a


def propagate_as(frames: np.ndarray, lens_shifted: np.ndarray) -> np.ndarray:
    """angular_spectrum (angular_spectrum.cu:33-63) on complex128 frames."""
    buf = np.ascontiguousarray(frames, np.complex128).copy()
    B, H, W = buf.shape
    lib().or_propagate_as_f64(_p(buf), B, W, H, _p(np.ascontiguousarray(lens_shifted, np.complex128)))
    return buf


def propagate_fresnel(frames: np.ndarray, chirp: np.ndarray) -> np.ndarray:
    """fresnel_transform (fresnel_transform.cu:65-83)."""
    buf = np.ascontiguousarray(frames, np.complex128).copy()
    B, H, W = buf.shape
    lib().or_propagate_fresnel_f64(_p(buf), B, W, H, _p(np.ascontiguousarray(chirp, np.complex128)))
    return buf


def stft(ring: np.ndarray) -> np.ndarray:
    """stft (stft.cu:9-15) over (T, H, W) in slot order."""
    ring = np.ascontiguousarray(ring, np.complex128)
    T = ring.shape[0]
    out = np.empty_like(ring)
    lib().or_stft_f64(_p(ring), T, ring[0].size, _p(out))
    return out


def image_type(spec: np.ndarray, img_type: int, pmin: int, pmax: int) -> np.ndarray:
    """complex_to_modulus/_squared_modulus/_argument (tools_conversion.cu)."""
    spec = np.ascontiguousarray(spec, np.complex128)
    out = np.empty(spec.shape[1:], np.float64)
    lib().or_image_type_f64(_p(spec), spec[0].size, img_type, pmin, pmax, _p(out))
    return out


def fft_freqs(T: int, fps: float) -> np.ndarray:
    out = np.empty(T, np.float32)
    lib().or_fft_freqs(T, fps, _p(out))
    return out


def moments(spec: np.ndarray, f_start: int, f_end: int, fps: float):
    spec = np.ascontiguousarray(spec, np.complex128)
    T = spec.shape[0]
    m = [np.empty(spec.shape[1:], np.float64) for _ in range(3)]
    lib().or_moments_f64(_p(spec), T, spec[0].size, f_start, f_end, fps, _p(m[0]), _p(m[1]), _p(m[2]))
    return m


def pca_cov(ring: np.ndarray) -> np.ndarray:
    """cov = H^H H (matrix_operations.cc:9-26), T x T, returned as a numpy
    matrix cov[i, j] (the column-major buffer read back)."""
    ring = np.ascontiguousarray(ring, np.complex128)
    T = ring.shape[0]
    cov = np.empty(T * T, np.complex128)
    lib().or_pca_cov_f64(_p(ring), T, ring[0].size, _p(cov))
    return cov.reshape(T, T).T.copy()  # column-major -> [i, j]


def eig_jacobi(cov: np.ndarray):
    """FP64 cyclic Jacobi; returns (evals ascending, V with V[:, j] the j-th
    eigenvector, sweeps)."""
    T = cov.shape[0]
    colmajor = np.ascontiguousarray(cov.T, np.complex128)  # [j, i] -> buffer i + j*T
    ev = np.empty(T, np.float64)
    vec = np.empty(T * T, np.complex128)
    sweeps = lib().or_eig_jacobi_f64(_p(colmajor), T, _p(ev), _p(vec))
    V = vec.reshape(T, T).T.copy()
    return ev, V, sweeps


def pca_project(ring: np.ndarray, V: np.ndarray) -> np.ndarray:
    """C = H V (matrix_operations.cc:77-102); returns (T, H, W)."""
    ring = np.ascontiguousarray(ring, np.complex128)
    T = ring.shape[0]
    colmajor = np.ascontiguousarray(V.T, np.complex128)
    out = np.empty_like(ring)
    lib().or_pca_project_f64(_p(ring), _p(colmajor), T, ring[0].size, _p(out))
    return out


def unwrap2d(phase: np.ndarray) -> np.ndarray:
    """unwrap_2d (tools_unwrap.cu:269-369), raw output before rescale."""
    phase = np.ascontiguousarray(phase, np.float64)
    H, W = phase.shape
    out = np.empty_like(phase)
    lib().or_unwrap2d_f64(_p(phase), W, H, _p(out))
    return out


def rescale_unwrap(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(img)
    lib().or_rescale_unwrap_f32(_p(img), img.size, _p(out))
    return out


def subsample(img: np.ndarray, step: int) -> np.ndarray:
    """subsample_frame (frame_reshape.cu:278-330)."""
    img = np.ascontiguousarray(img)
    H, W = img.shape
    out = np.empty((H // step, W // step), img.dtype)
    if img.dtype == np.float32:
        lib().or_subsample_f32(_p(img), W, H, step, _p(out))
    elif img.dtype == np.uint8:
        lib().or_subsample_u8(_p(img), W, H, step, _p(out))
    else:
        raise TypeError(img.dtype)
    return out


def percentile(img: np.ndarray, pcts) -> np.ndarray:
    """compute_percentile (percentile.cu:75-94), float index rule."""
    img = np.ascontiguousarray(img, np.float32).ravel()
    p = np.asarray(pcts, np.float32)
    out = np.empty(len(p), np.float32)
    lib().or_percentile_f32(_p(img), img.size, _p(p), len(p), _p(out))
    return out


def reticle_pixels(img: np.ndarray, scale: float) -> np.ndarray:
    """The pixels of the contrast reticle's centred disc (percentile.cu:15-43,
    150-190), row-major order."""
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape
    out = np.empty(img.size, np.float32)
    n = lib().or_reticle_pixels_f32(_p(img), W, H, scale, _p(out))
    return out[:n].copy()


def apply_contrast(img: np.ndarray, mn: float, mx: float) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32).copy()
    lib().or_apply_contrast_f32(_p(img), img.size, mn, mx)
    return img


def float_to_u16(img: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    out = np.empty(img.shape, np.uint16)
    lib().or_float_to_u16(_p(img), img.size, _p(out))
    return out


# --------------------------------------------------------------------------
# Pipe
# --------------------------------------------------------------------------
class OraclePipe:
    """Per-batch Pipe cycle (pipe.cc:295-455) on the CPU.  `settings` is a
    ctypes hv_compute_settings structure (holovibes_amd.settings layout)."""

    def __init__(self, settings, precision: int = 64):
        self.precision = precision
        self.settings = settings
        self.P = settings.width * settings.height
        self.h = lib().or_pipe_create(ctypes.byref(settings), precision)

    def push(self, frames: np.ndarray):
        """Returns None when the cycle exited early, else a dict with the
        float image, the contrasted float image, the u16 output and the
        contrast range."""
        frames = np.ascontiguousarray(frames)
        s = self.settings
        image = np.empty((s.height, s.width), np.float32)
        cimg = np.empty_like(image)
        out = np.empty((s.height, s.width), np.uint16)
        rng = np.empty(2, np.float32)
        r = lib().or_pipe_push(self.h, self.precision, _p(frames), _p(image), _p(cimg), _p(out), _p(rng))
        if not r:
            return None
        return {"image": image, "contrast": cimg, "u16": out, "range": rng}

    def ring_state(self):
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        lib().or_pipe_ring_state(self.h, self.precision, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def enqueue_raw(self, frames: np.ndarray):
        frames = np.ascontiguousarray(frames, np.complex128)
        lib().or_pipe_enqueue_raw(self.h, _p(frames), frames.shape[0])

    def ring(self) -> np.ndarray:
        s = self.settings
        out = np.empty((s.time_transformation_size, s.height, s.width), np.complex128)
        lib().or_pipe_ring_copy(self.h, _p(out))
        return out

    def moments(self):
        """M0, M1, M2 (float64, (H, W)) of the latest cycle's time transform over
        the settings' range [p_start, min(p_start + p_width, T - 1)]."""
        s = self.settings
        m = [np.empty((s.height, s.width), np.float64) for _ in range(3)]
        lib().or_pipe_moments(self.h, self.precision, *[_p(x) for x in m])
        return m

    def pca_state(self):
        """(cov[i, j], eigenvalues ascending, V with V[:, j] the j-th
        eigenvector) of the latest PCA cycle, complex128 / float64."""
        T = self.settings.time_transformation_size
        cov = np.empty(T * T, np.complex128)
        ev = np.empty(T, np.float64)
        vec = np.empty(T * T, np.complex128)
        lib().or_pipe_pca_state(self.h, self.precision, _p(cov), _p(ev), _p(vec))
        return cov.reshape(T, T).T.copy(), ev, vec.reshape(T, T).T.copy()

    def close(self):
        if self.h:
            lib().or_pipe_destroy(self.h, self.precision)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
