"""CPU restatement of the optional post-processing of the float image
(test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the pipe's
renormalisation, image accumulation and log10; the product package never
imports it.  float32 arithmetic in the reference's operation order.
"""
from __future__ import annotations

import numpy as np


def renormalize(x: np.ndarray, norm_constant: int) -> np.ndarray:
    """gpu_normalize (Backend/cuda_sources/tools_compute.cu:99-123): the sum
    in double, then x * 2^c * (frame_res / (float)sum), left to right in float."""
    x = np.asarray(x, np.float32)
    total = np.float32(np.sum(x, dtype=np.float64))
    ratio = np.float32(x.size) / total
    return (x * np.float32(1 << norm_constant)) * ratio


class Accumulator:
    """Queue of the last `level` images (Backend/sources/data_struct/queue.cc:149-189)
    and their mean summed from the oldest in float (kernel_accumulate_images,
    Backend/cuda_sources/tools_conversion.cu:418-441)."""

    def __init__(self, level: int):
        self.level, self.images = level, []

    def push(self, img: np.ndarray) -> np.ndarray:
        self.images.append(np.asarray(img, np.float32).copy())
        self.images = self.images[-self.level:]
        acc = np.zeros_like(self.images[0])
        for im in self.images:
            acc = acc + im
        return acc / np.float32(len(self.images))


def log10(x: np.ndarray) -> np.ndarray:
    """map_log10 (Backend/cuda_sources/map.cu:3-8), correctly rounded: log10
    in double, rounded once to float (CUDA's log10f is within 2 ulp of it)."""
    return np.log10(np.asarray(x, np.float32).astype(np.float64)).astype(np.float32)


def shift_corners(x: np.ndarray) -> np.ndarray:
    """Quadrant swap of an even-sided image (Backend/cuda_sources/shift_corners.cu:8-44)."""
    return np.fft.fftshift(np.asarray(x, np.float32))


def cut_values(z: np.ndarray, img_type: int) -> np.ndarray:
    """The per-pixel value of fill_32bit_slices (Backend/cuda_sources/stft.cu:36-49):
    modulus, its square, atan(im / re) + pi/2, or 0 (other image types);
    correctly rounded where the reference calls a float math function."""
    z = np.asarray(z, np.complex64)
    re, im = z.real.astype(np.float32), z.imag.astype(np.float32)
    if img_type == 0:
        return np.hypot(re.astype(np.float64), im.astype(np.float64)).astype(np.float32)
    if img_type == 1:
        m = np.hypot(re.astype(np.float64), im.astype(np.float64)).astype(np.float32)
        return m * m
    if img_type == 2:
        with np.errstate(divide="ignore", invalid="ignore"):
            q = im / re
        a = np.arctan(q.astype(np.float64)).astype(np.float32)
        return (a.astype(np.float64) + np.pi / 2).astype(np.float32)
    return np.zeros(z.shape, np.float32)


def cuts(stft: np.ndarray, xmin: int, ymin: int, xmax: int, ymax: int, img_type: int):
    """time_transformation_cuts_begin (stft.cu:75-115) of a (T, H, W) window:
    XZ (T, W) = mean over rows ymin..ymax, YZ (H, T) = mean over columns
    xmin..xmax, summed in float in index order."""
    v = cut_values(stft, img_type)
    T, H, W = v.shape
    xz = np.zeros((T, W), np.float32)
    for y in range(ymin, ymax + 1):
        xz = xz + v[:, y, :]
    xz = xz / np.float32(ymax - ymin + 1)
    yz = np.zeros((H, T), np.float32)
    for x in range(xmin, xmax + 1):
        yz = yz + v[:, :, x].T
    yz = yz / np.float32(xmax - xmin + 1)
    return xz, yz
