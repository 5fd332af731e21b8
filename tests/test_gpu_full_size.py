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


@pytest.mark.parametrize("level", [2, 0])
@pytest.mark.parametrize("precision,tol", [(0, 1e-5), (1, 1e-4), (2, 1e-4)])
def test_c2_full_size(oracle, precision, tol, level):
    """c2: 512^2 u8, Fresnel, PCA T = 32 (FP32 / split BF16 / split FP16
    covariance, FP64 Jacobi), squared modulus of eigenvectors 4..31, on the
    bins-first path (2) and in the reference's step order (0)."""
    print(f"c2 512x512 T=32 pca_precision={precision} path {level}", flush=True)
    s, fl = frames("c2", 2, pca_precision=precision)
    check(run_both(oracle, s, fl, fused=level, path=level), tol)


def test_c3_full_size(oracle):
    """c3: 1024^2 u16, T = 64, B = S = 32 (the ring keeps half of the window
    and rotates: physical slot order), argument of bin 2, 2-D unwrap,
    percentile 0.5 / 99.5 % on an 8 x 8 subsample; three cycles."""
    print("c3 1024x1024 T=64 B=32", flush=True)
    s, fl = frames("c3", 4)
    res = run_both(oracle, s, fl)
    assert len(res) == 3
    check(res, 1e-3)


def test_c5_stft_full_size(oracle):
    """c5 STFT variant: 2048^2 u16, B = S = T = 128, angular spectrum,
    moment 0 and moment 1 over bins 4..60 (camera fps 128), on the
    bins-first path (57 planes propagated) and on the three-pass path (128
    frames propagated), against one FP64 oracle run."""
    from holovibes_amd.constants import HV_IMG_MOMENTS_0
    from holovibes_amd.pipe import Pipe
    s, fl = frames("c5_stft", 1)
    print("c5_stft 2048x2048 T=128: FP64 oracle", flush=True)
    o = oracle.OraclePipe(s, 64)
    ref = o.push(fl[0])
    m_ref = o.moments()
    o.close()
    assert rel_l2(ref["image"], m_ref[0]) < 1e-6
    dev = torch.from_numpy(fl[0].view(np.int16)).cuda()
    torch.cuda.synchronize()
    for level in (2, 0):
        for k in (0, 1):
            print(f"c5_stft path {level} moment {k}", flush=True)
            s2 = type(s).from_buffer_copy(s)
            s2.img_type = HV_IMG_MOMENTS_0 + k
            p = Pipe(s2, fused=level)
            assert p.path == level
            assert p.push(dev)
            e = rel_l2(p.image().cpu().numpy(), m_ref[k])
            p.close()
            print(f"  image rel L2 vs FP64 {e:.3g}", flush=True)
            assert e < 1e-5, (level, k, e)


def test_c5_pca_full_size(oracle):
    """c5 PCA variant: 2048^2 u16, B = S = T = 128, angular spectrum, PCA
    with the covariance in plain FP32, split BF16 and split FP16 (the
    configured one), FP64 Jacobi; moment 0 and moment 1 of components 4..60
    on the bins-first path (covariance of the raw window, 57 components
    propagated) and in the reference's step order, against the FP64 oracle
    at 1e-4 relative L2.  The window's spectrum has
    13 near-degenerate noise eigenvalues 1e-9 lambda_max apart and the range
    starts among them, so this bar needs the covariance centred and
    assembled in FP64.  The covariance (step function, float output) and the
    eigenvalues are compared with the oracle's FP64 ones too."""
    from holovibes_amd import steps as st
    from holovibes_amd.constants import HV_IMG_MOMENTS_0
    from holovibes_amd.pipe import Pipe
    s, fl = frames("c5_pca", 1)
    print("c5_pca 2048x2048 T=128: FP64 oracle", flush=True)
    o = oracle.OraclePipe(s, 64)
    assert o.push(fl[0]) is not None
    m_ref = o.moments()
    cov64, ev64, _ = o.pca_state()
    o.close()
    gaps = np.diff(ev64) / ev64[-1]
    print(f"eigen-gaps / lambda_max at the range edges: 3|4 {gaps[3]:.3g}, 60|61 {gaps[60]:.3g}", flush=True)
    dev = torch.from_numpy(fl[0].view(np.int16)).cuda()
    torch.cuda.synchronize()
    errs = {}
    for precision in (0, 1, 2):
        for level in (2, 0):
            for k in (0, 1):
                print(f"c5_pca pca_precision={precision} path {level} moment {k}", flush=True)
                s2 = type(s).from_buffer_copy(s)
                s2.pca_precision = precision
                s2.img_type = HV_IMG_MOMENTS_0 + k
                p = Pipe(s2, fused=level)
                assert p.path == level
                assert p.push(dev)
                errs[(precision, level, k)] = rel_l2(p.image().cpu().numpy(), m_ref[k])
                if level == 0 and k == 0:  # the three-pass ring holds the propagated window
                    ring, _, _ = p.ring()
                    cov = st.pca_cov(ring, precision).cpu().numpy()
                    ecov = rel_l2(cov, cov64)
                    ev, _ = st.eig_hermitian(torch.from_numpy(cov).cuda())
                    eev = float(np.abs(ev.cpu().numpy() - ev64).max() / ev64[-1])
                    print(f"  covariance rel L2 {ecov:.3g}, eigenvalues max err / lambda_max {eev:.3g}", flush=True)
                    assert ecov < 2e-6 and eev < 1e-6, (precision, ecov, eev)
                    del ring
                p.close()
                print(f"  image rel L2 vs FP64 {errs[(precision, level, k)]:.3g}", flush=True)
    for key, e in errs.items():
        assert e < 1e-4, (key, e)
