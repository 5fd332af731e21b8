"""Frame sides that are not powers of two (camera geometries 2^a 3^b 5^c such
as 1536 x 2048, 1920 x 1080) and time windows that are not (T = 24, ...): the
runtime-length FFT and time-DFT kernels against the oracle's mixed-radix FFT
(pinned by numpy.fft at these lengths), the padded eigensolver against
numpy's eigh, and whole pipes against the oracle's pipe on the same seeded
frames.

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


# ---------------------------------------------------------------------------
# time windows that are not powers of two
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("T", [3, 12, 24, 100])
def test_stft_other_windows(st, oracle, T):
    rng = np.random.default_rng(14)
    ring = crand(rng, T, 16, 32)
    got = host(st.stft(torch.from_numpy(ring).cuda()))
    assert rel_l2(got, oracle.stft(ring.astype(np.complex128))) < 1e-6


@pytest.mark.parametrize("T,pmin,pmax", [(24, 2, 7), (12, 0, 11), (100, 5, 5), (15, 3, 14)])
@pytest.mark.parametrize("img_type", [0, 1, 2])
def test_stft_image_other_windows(st, oracle, T, pmin, pmax, img_type):
    rng = np.random.default_rng(15)
    ring = crand(rng, T, 32, 32) * 100
    img, mm = st.stft_image(torch.from_numpy(ring).cuda(), img_type, pmin, pmax, minmax=True)
    got = host(img)
    want = oracle.image_type(oracle.stft(ring.astype(np.complex128)), img_type, pmin, pmax)
    assert rel_l2(got, want) < (2e-6 if img_type != 2 else 2e-5)
    r = host(st.decode_minmax(mm))
    assert r[0] == got.min() and r[1] == got.max()


@pytest.mark.parametrize("T,fs,fe", [(24, 2, 20), (24, 0, 23), (15, 1, 13)])
def test_stft_moments_other_windows(st, oracle, T, fs, fe):
    rng = np.random.default_rng(16)
    ring = crand(rng, T, 16, 64)
    fps = 96.0
    m = [host(x) for x in st.stft_moments(torch.from_numpy(ring).cuda(), fs, fe, fps)]
    w = oracle.moments(oracle.stft(ring.astype(np.complex128)), fs, fe, fps)
    for k in range(3):
        scale = np.abs(w[0]).max() * (fps if k else 1) ** k
        assert np.abs(m[k] - w[k]).max() < 2e-5 * scale


@pytest.mark.parametrize("T", [3, 12, 24, 31])
def test_eig_hermitian_other_windows(st, oracle, T):
    """The eigensolver pads T to the next power of two with a decoupled block
    above every eigenvalue; the T smallest eigenpairs are the matrix's."""
    rng = np.random.default_rng(17)
    A = crand(rng, T, 3 * T).astype(np.complex128)
    A = (A @ A.conj().T).astype(np.complex64)
    ev, V = st.eig_hermitian(torch.from_numpy(A).cuda())
    ev, V = host(ev), host(V)
    assert ev.shape == (T,) and V.shape == (T, T)
    A64 = A.astype(np.complex128)
    w = np.linalg.eigvalsh(A64)
    assert np.allclose(ev, w, rtol=1e-10, atol=1e-10 * np.abs(w).max())
    wo, Vo, _ = oracle.eig_jacobi(A64)
    assert np.abs(V - Vo).max() < 1e-8
    assert rel_l2(A64 @ V, V * ev[None, :]) < 1e-12


def test_pca_image_other_window(st, oracle):
    rng = np.random.default_rng(18)
    T = 24
    ring = crand(rng, T, 32, 32) * 50
    V = np.linalg.qr(rng.standard_normal((T, T)) + 1j * rng.standard_normal((T, T)))[0]
    got = host(st.pca_image(torch.from_numpy(ring).cuda(), torch.from_numpy(V.astype(np.complex128)).cuda(), 1, 4, 23))
    want = oracle.image_type(oracle.pca_project(ring.astype(np.complex128), V), 1, 4, 23) * 131071
    assert rel_l2(got, want) < 2e-6


@pytest.mark.parametrize("img_type", [0, 2])
def test_pipe_window_24(oracle, img_type):
    """B = S = T = 24 (three-pass path: the bins-first and fused kernels are
    built for power-of-two windows)."""
    s, fl = frames("c1", 2, width=64, height=64, batch_size=24, time_stride=24, time_transformation_size=24,
                   img_type=img_type, p_start=2, p_width=5)
    check(run_both(oracle, s, fl, path=0), 1e-5 if img_type == 0 else 1e-4)


def test_pipe_window_24_rotation(oracle):
    """T = 24, B = S = 8: the window rotates in physical slot order (n mod 24)."""
    s, fl = frames("c1", 5, width=64, height=64, batch_size=8, time_stride=8, time_transformation_size=24,
                   p_start=1, p_width=4)
    res = run_both(oracle, s, fl)
    assert len(res) == 3
    check(res, 1e-5)


def test_pipe_window_24_moments(oracle):
    from holovibes_amd.constants import HV_IMG_MOMENTS_1
    s, fl = frames("c5_stft", 1, width=64, height=64, batch_size=24, time_stride=24, time_transformation_size=24,
                   p_start=2, p_width=18, camera_fps=24.0, img_type=HV_IMG_MOMENTS_1)
    check(run_both(oracle, s, fl), 1e-5)


def test_pipe_window_24_pca(oracle):
    """PCA over a 24-frame window: MFMA covariance of one frame group, the
    padded eigensolver, projection onto eigenvectors 4..23."""
    s, fl = frames("c2", 1, width=64, height=64, batch_size=24, time_stride=24, time_transformation_size=24,
                   p_start=4, p_width=19)
    check(run_both(oracle, s, fl), 1e-4)


def test_pipe_camera_frame_window_24(oracle):
    """1536 x 2048 camera frames with T = 24 (c1's other settings)."""
    s, fl = frames("c1", 1, width=1536, height=2048, batch_size=24, time_stride=24, time_transformation_size=24)
    check(run_both(oracle, s, fl), 1e-5)
