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
