"""CPU restatement of the optional post-processing of the float image
(test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the pipe's
renormalisation, image accumulation, log10, registration and XZ/YZ cuts;
the product package never imports it.  float32 arithmetic in the
reference's operation order.
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


def registration_mask(W: int, H: int, zone: float, swap: bool) -> np.ndarray:
    """get_circular_mask (Backend/cuda_sources/masks.cu:61-77): 1 inside the
    disc of radius (min(W, H) / 2) * zone around (W/2, H/2), quadrant-swapped
    when exactly one of Fresnel and the fft shift holds (registration.cc:100-110)."""
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    r = np.float32(min(W, H) // 2) * np.float32(zone)
    m = (((x - np.float32(W / 2)) ** 2 + (y - np.float32(H / 2)) ** 2) <= r * r).astype(np.float32)
    return np.fft.fftshift(m) if swap else m


def registration_prepare(img: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """image_preprocess (registration.cc:58-63): the mean inside the mask
    removed inside it, 0 outside."""
    img = np.asarray(img, np.float32)
    count = float(mask.sum(dtype=np.float64))
    mean = np.float32(np.float32((img * mask).sum(dtype=np.float64)) / np.float32(count)) if count > 0 else np.float32(0)
    return (img - mean) * mask


def register(img: np.ndarray, ref: np.ndarray, mask: np.ndarray):
    """One registration (registration.cc:11-56): the peak (first largest
    |value|) of the cross-correlation of the prepared images, wrapped to a
    signed shift and negated; the image shifted circularly by it.  Returns
    (shifted image, (shift_x, shift_y))."""
    a = np.fft.fft2(registration_prepare(img, mask).astype(np.float64))
    b = np.fft.fft2(registration_prepare(ref, mask).astype(np.float64))
    xc = np.fft.ifft2(a * np.conj(b)).real
    H, W = img.shape
    idx = int(np.argmax(np.abs(xc)))
    x, y = idx % W, idx // W
    if x > W // 2:
        x -= W
    if y > H // 2:
        y -= H
    sx, sy = -x, -y
    return np.roll(np.asarray(img, np.float32), (sy, sx), axis=(0, 1)), (sx, sy)
