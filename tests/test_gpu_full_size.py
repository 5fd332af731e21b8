"""Every workload at its full size on the GPU pipe against the FP64 oracle
pipe (pipe.cc:295-455 restated) on the same seeded frames.

These instantiate the kernels that only full sizes reach: the 2048-long
column pass (4-column tiles), the T = 128 STFT in LDS over a 4 GiB ring, the
grouped half-precision covariance at 4.2 M pixels, the 1024^2 unwrap and the
8 x 8 subsampled percentile, the 512^2 Fresnel + PCA at each covariance
precision.  Bars as in test_gpu_pipe.py: float image within relative L2
1e-5 of FP64 (split-precision PCA 1e-4; the unwrapped phase, rescaled by its
own extremes, 1e-3), u16 within one count and identical on >= 95 % of the
pixels.
"""
import numpy as np
import pytest
import torch

from test_gpu_pipe import check, frames, rel_l2, run_both

pytestmark = pytest.mark.gpu


def test_c1_full_size(oracle):
    """c1: 256^2 u16, B = S = T = 16, angular spectrum, STFT bins 2..7, min-max."""
    print("c1 256x256 T=16", flush=True)
    s, fl = frames("c1", 3)
    check(run_both(oracle, s, fl), 1e-5)


@pytest.mark.parametrize("precision,tol", [(0, 1e-5), (1, 1e-4), (2, 1e-4)])
def test_c2_full_size(oracle, precision, tol):
    """c2: 512^2 u8, Fresnel, PCA T = 32 (FP32 / split BF16 / split FP16
    covariance, FP64 Jacobi), squared modulus of eigenvectors 4..31."""
    print(f"c2 512x512 T=32 pca_precision={precision}", flush=True)
    s, fl = frames("c2", 2, pca_precision=precision)
    check(run_both(oracle, s, fl), tol)


def test_c3_full_size(oracle):
    """c3: 1024^2 u16, T = 64, B = S = 32 (the ring keeps half of the window
    and rotates: physical slot order), argument of bin 2, 2-D unwrap,
    percentile 0.5 / 99.5 % on an 8 x 8 subsample; three cycles."""
    print("c3 1024x1024 T=64 B=32", flush=True)
    s, fl = frames("c3", 4)
    res = run_both(oracle, s, fl)
    assert len(res) == 3
    check(res, 1e-3)


@pytest.mark.parametrize("img_type", [5, 6])
def test_c5_stft_full_size(oracle, img_type):
    """c5 STFT variant: 2048^2 u16, B = S = T = 128, angular spectrum,
    moment 0 and moment 1 over bins 4..60 (camera fps 128)."""
    print(f"c5_stft 2048x2048 T=128 moment {img_type - 5}", flush=True)
    s, fl = frames("c5_stft", 1, img_type=img_type)
    check(run_both(oracle, s, fl), 1e-5)
