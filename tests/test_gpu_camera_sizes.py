"""Frame sides that are not powers of two (camera geometries 2^a 3^b 5^c such
as 1536 x 2048, 1920 x 1080): the runtime-length FFT kernels against the
oracle's mixed-radix FFT (pinned by numpy.fft at these lengths), and whole
pipes against the oracle's pipe on the same seeded frames.

cufftPlanMany accepts any n (Backend/sources/core/icompute.cc:109-115,
151-175); the reference's path itself is the same for every size.

Bars: FP32 FFTs within relative L2 1e-6 of the FP64 oracle; float images
within 1e-5 (PCA split-BF16 1e-4); u16 within 1 count on every pixel and
identical on >= 95 %.
"""
import numpy as np
import pytest
import torch

from test_gpu_pipe import check, frames, run_both

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def st():
    from holovibes_amd import steps
    steps._lib.load()
    assert torch.cuda.is_available()
    return steps


def rel_l2(a, b):
    a, b = np.asarray(a, np.complex128), np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def crand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# (batch, H, W): both sides generic, one side generic (the other on the
# templated kernels), a width that is no multiple of the column tile under a
# power-of-two height, columns longer than 2048 (two-column tiles), 1080p
FFT_SHAPES = [(2, 24, 40), (1, 96, 160), (1, 48, 1536), (1, 1536, 48), (1, 64, 1350), (1, 3000, 16),
              (1, 1080, 1920), (1, 2048, 1536), (1, 1536, 2048)]


@pytest.mark.parametrize("shape", FFT_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fft2_camera_sizes(st, oracle, shape):
    rng = np.random.default_rng(11)
    x = crand(rng, *shape)
    for d in (-1, 1):
        t = torch.from_numpy(x).cuda()
        st.fft2(t, d)
        err = rel_l2(host(t), oracle.fft2(x.astype(np.complex128), d))
        assert err < 1e-6, (shape, d, err)


@pytest.mark.parametrize("shape", [(2, 96, 160), (1, 1080, 1920)], ids=lambda s: "x".join(map(str, s)))
def test_angular_spectrum_camera_sizes(st, oracle, shape):
    rng = np.random.default_rng(12)
    B, H, W = shape
    x = crand(rng, *shape) * 1000
    lens = oracle.lens_as(W, H, 0.25, 852e-9, 20.0).astype(np.complex64)
    t = torch.from_numpy(x).cuda()
    st.angular_spectrum(t, torch.from_numpy(lens).cuda())
    want = oracle.propagate_as(x.astype(np.complex128), lens.astype(np.complex128))
    assert rel_l2(host(t), want) < 2e-6


def test_fresnel_camera_size(st, oracle):
    rng = np.random.default_rng(13)
    x = crand(rng, 2, 120, 96)
    c = oracle.lens_fresnel(96, 120, 785e-9, 0.4, 12.0).astype(np.complex64)
    t = torch.from_numpy(x).cuda()
    st.fresnel_transform(t, torch.from_numpy(c).cuda())
    want = oracle.propagate_fresnel(x.astype(np.complex128), c.astype(np.complex128))
    assert rel_l2(host(t), want) < 2e-6


@pytest.mark.parametrize("level", [2, 0])
@pytest.mark.parametrize("W,H", [(192, 160), (160, 96), (96, 240)])
def test_pipe_camera_sizes(oracle, W, H, level):
    """B = S = T angular spectrum + STFT on the bins-first split path (2) and
    the three-pass path (0)."""
    s, fl = frames("c1", 2, width=W, height=H)
    check(run_both(oracle, s, fl, fused=level, path=level), 1e-5)


def test_pipe_camera_size_fresnel_pca(oracle):
    from holovibes_amd.constants import HV_DEPTH_8
    s, fl = frames("c2", 1, width=120, height=96, depth=HV_DEPTH_8)
    check(run_both(oracle, s, fl), 1e-4)


def test_pipe_camera_size_ring_rotation_unwrap(oracle):
    """T = 2B (the ring rotates), argument + 2-D unwrap + percentile on an
    8 x 8 subsample (the c3 configuration) at 96 x 80."""
    s, fl = frames("c3", 4, width=96, height=80, batch_size=8, time_stride=8, time_transformation_size=16)
    res = run_both(oracle, s, fl)
    assert len(res) == 3
    check(res, 1e-4)


@pytest.mark.parametrize("W,H", [(2048, 1536), (1920, 1080)])
def test_pipe_camera_full_frame(oracle, W, H):
    """A full camera frame (c1's settings: B = S = T = 16, angular spectrum,
    STFT bins 2..7), bins-first and three-pass against one oracle run."""
    from holovibes_amd.pipe import Pipe
    s, fl = frames("c1", 1, width=W, height=H)
    ref = oracle.OraclePipe(s, 64).push(fl[0])
    for level in (2, 0):
        p = Pipe(s, fused=level)
        assert p.path == level
        assert p.push(torch.from_numpy(fl[0].view(np.int16)).cuda())
        img, u = p.image().cpu().numpy(), p.output_u16()
        p.close()
        check([(img, u, None, ref)], 1e-5)
