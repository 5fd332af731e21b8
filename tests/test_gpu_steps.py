"""GPU parity of each step function (libholo_hip.so through its C-ABI) against
the pinned CPU oracle on the same seeded inputs.

Bars: integer / index work bit-exact; FP32 transforms within relative L2
1e-6 of the FP64 oracle per FFT pass (stated per test); reductions exact.
"""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))


@pytest.fixture(scope="module")
def st():
    from holovibes_amd import steps
    steps._lib.load()
    assert torch.cuda.is_available()
    return steps


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def crand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# ---------------------------------------------------------------------------
def test_convert_bit_exact(st, oracle):
    rng = np.random.default_rng(1)
    for dt in (np.uint8, np.uint16):
        raw = rng.integers(0, np.iinfo(dt).max + 1, (3, 16, 32)).astype(dt)
        t = dev(raw.view(np.int16) if dt == np.uint16 else raw)
        got = host(st.convert(t))
        assert np.array_equal(got, oracle.convert(raw).astype(np.complex64))
    f = rng.standard_normal((2, 16, 16)).astype(np.float32) * 1e4
    assert np.array_equal(host(st.convert(dev(f))), f.astype(np.complex64))


@pytest.mark.parametrize("precision", [0, 1])
def test_lens_angular_spectrum(st, oracle, precision):
    for nx, ny, z, px in [(64, 32, 0.25, 20.0), (512, 512, 0.25, 20.0), (256, 128, 0.3, 0.3)]:
        got = host(st.lens_angular_spectrum(nx, ny, z, 852e-9, px, precision))
        want = oracle.lens_as(nx, ny, z, 852e-9, px, precision, shifted=False)
        assert np.abs(got - want).max() < 2e-6, (nx, ny)


@pytest.mark.parametrize("precision", [0, 1])
def test_lens_fresnel(st, oracle, precision):
    for W, H in [(64, 64), (512, 512), (64, 32), (32, 64)]:
        got = host(st.lens_fresnel(W, H, 785e-9, 0.4, 12.0, precision))
        want = oracle.lens_fresnel(W, H, 785e-9, 0.4, 12.0, precision)
        assert np.abs(got - want).max() < 2e-6, (W, H)


def test_prepare_lens_is_shifted_scaled(st, oracle):
    from holovibes_amd.constants import HV_SPACE_ANGULARSP
    got = host(st.prepare_lens(HV_SPACE_ANGULARSP, 128, 64, 852e-9, 0.25, 20.0))
    want = oracle.lens_as(128, 64, 0.25, 852e-9, 20.0) / (128 * 64)
    assert np.abs(got - want).max() < 2e-6 / (128 * 64)


def test_shift_corners_exact(st):
    rng = np.random.default_rng(2)
    a = crand(rng, 3, 32, 64)
    t = dev(a)
    st.shift_corners(t)
    assert np.array_equal(host(t), np.fft.fftshift(a, axes=(-2, -1)))
    f = rng.standard_normal((16, 8)).astype(np.float32)
    tf = dev(f)
    st.shift_corners(tf)
    assert np.array_equal(host(tf), np.fft.fftshift(f))


FFT_SHAPES = [(2, 16, 16), (1, 16, 4096), (1, 4096, 16), (2, 32, 64), (3, 64, 128), (1, 256, 256), (2, 512, 512),
              (1, 1024, 1024), (1, 2048, 32), (1, 32, 2048),
              # long columns in tiles narrower than a cache line, with the XCD-aware tile order active
              (1, 2048, 128), (1, 4096, 128)]


@pytest.mark.parametrize("shape", FFT_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_fft2(st, oracle, shape):
    rng = np.random.default_rng(3)
    x = crand(rng, *shape)
    for d in (-1, 1):
        t = dev(x)
        st.fft2(t, d)
        err = rel_l2(host(t), oracle.fft2(x.astype(np.complex128), d))
        assert err < 1e-6, (shape, d, err)


@pytest.mark.parametrize("shape", [(4, 64, 64), (2, 512, 512), (1, 256, 1024), (2, 1024, 128), (1, 2048, 128),
                                   (1, 4096, 128)], ids=lambda s: "x".join(map(str, s)))
def test_angular_spectrum(st, oracle, shape):
    rng = np.random.default_rng(4)
    B, H, W = shape
    x = crand(rng, *shape) * 1000
    lens = oracle.lens_as(W, H, 0.25, 852e-9, 20.0)
    t = dev(x)
    st.angular_spectrum(t, dev(lens.astype(np.complex64)))
    want = oracle.propagate_as(x.astype(np.complex128), lens.astype(np.complex64).astype(np.complex128))
    assert rel_l2(host(t), want) < 2e-6


def test_fresnel_transform(st, oracle):
    rng = np.random.default_rng(5)
    x = crand(rng, 3, 64, 128)
    c = oracle.lens_fresnel(128, 64, 785e-9, 0.4, 12.0).astype(np.complex64)
    t = dev(x)
    st.fresnel_transform(t, dev(c))
    want = oracle.propagate_fresnel(x.astype(np.complex128), c.astype(np.complex128))
    assert rel_l2(host(t), want) < 2e-6


@pytest.mark.parametrize("space", [0, 1, 2])
@pytest.mark.parametrize("B,T,slot0", [(5, 8, 6), (4, 4, 0), (10, 4, 1)])
def test_propagate_to_ring(st, oracle, space, B, T, slot0):
    """Fused conversion + space transform + queue transfer: frame b lands in
    slot (slot0 + b) mod T; of an oversized batch only the last T frames."""
    rng = np.random.default_rng(6)
    H, W = 64, 32
    raw = rng.integers(0, 4096, (B, H, W)).astype(np.uint16)
    lens = st.prepare_lens(space, W, H, 852e-9, 0.25 if space == 2 else 0.4, 20.0) if space else None
    init = crand(rng, T, H, W)
    ring = dev(init)
    st.propagate_to_ring(dev(raw.view(np.int16)), space, lens, ring, slot0)
    got = host(ring)
    conv = oracle.convert(raw)
    if space == 2:
        prop = oracle.propagate_as(conv, oracle.lens_as(W, H, 0.25, 852e-9, 20.0))
    elif space == 1:
        prop = oracle.propagate_fresnel(conv, oracle.lens_fresnel(W, H, 852e-9, 0.4, 20.0))
    else:
        prop = conv
    want = init.astype(np.complex128).copy()
    for b in range(B):
        want[(slot0 + b) % T] = prop[b]
    assert rel_l2(got, want) < 2e-6


@pytest.mark.parametrize("T", [1, 2, 4, 8, 16, 32, 64, 128, 256])
def test_stft(st, oracle, T):
    rng = np.random.default_rng(7)
    ring = crand(rng, T, 16, 32)
    got = host(st.stft(dev(ring)))
    assert rel_l2(got, oracle.stft(ring.astype(np.complex128))) < 1e-6


@pytest.mark.parametrize("T,pmin,pmax", [(1, 0, 0), (8, 0, 7), (16, 2, 7), (32, 2, 15), (64, 5, 5), (128, 4, 60),
                                         (256, 0, 255),
                                         # <= 4 bins: direct accumulation instead of the FFT
                                         (16, 0, 3), (64, 2, 2), (128, 61, 64), (256, 250, 253)])
@pytest.mark.parametrize("img_type", [0, 1, 2])
def test_stft_image(st, oracle, T, pmin, pmax, img_type):
    rng = np.random.default_rng(8)
    ring = crand(rng, T, 32, 32) * 100
    img, mm = st.stft_image(dev(ring), img_type, pmin, pmax, minmax=True)
    got = host(img)
    want = oracle.image_type(oracle.stft(ring.astype(np.complex128)), img_type, pmin, pmax)
    tol = 2e-6 if img_type != 2 else 2e-5  # the argument of near-zero bins is ill-conditioned
    assert rel_l2(got, want) < tol
    r = host(st.decode_minmax(mm))
    assert r[0] == got.min() and r[1] == got.max()
    # same result as the unfused reference sequence (stft + complex_to_*)
    spec = st.stft(dev(ring))
    assert np.allclose(host(st.complex_to_image(spec, img_type, pmin, pmax)), got, rtol=1e-5, atol=1e-5 * np.abs(got).max())


@pytest.mark.parametrize("T,fs,fe", [(16, 2, 7), (32, 0, 31), (128, 4, 60), (8, 4, 4)])
def test_stft_moments(st, oracle, T, fs, fe):
    rng = np.random.default_rng(9)
    ring = crand(rng, T, 16, 64)
    fps = 128.0
    m = [host(x) for x in st.stft_moments(dev(ring), fs, fe, fps)]
    w = oracle.moments(oracle.stft(ring.astype(np.complex128)), fs, fe, fps)
    for k in range(3):
        scale = np.abs(w[0]).max() * (fps if k else 1) ** k
        assert np.abs(m[k] - w[k]).max() < 2e-5 * scale


@pytest.mark.parametrize("precision,tol", [(0, 2e-6), (1, 2e-5), (2, 2e-6), (3, 1e-2), (4, 2e-3)])
@pytest.mark.parametrize("T,H,W", [(8, 64, 64), (16, 64, 64), (32, 64, 64), (64, 64, 64), (128, 64, 64),
                                   (256, 32, 64), (128, 256, 512)])
def test_pca_cov(st, oracle, T, H, W, precision, tol):
    """(T >= 64 in half precision runs the grouped kernel: every pair per
    workgroup; 128 x 256 x 512 gives it 512 slices of 256 pixels.)"""
    rng = np.random.default_rng(10)
    ring = crand(rng, T, H, W) * 300
    got = host(st.pca_cov(dev(ring), precision))
    want = oracle.pca_cov(ring.astype(np.complex128))
    err = rel_l2(got, want)
    assert err < tol, err
    assert np.allclose(got, got.conj().T, atol=1e-5 * np.abs(want).max())


@pytest.mark.parametrize("T", [2, 8, 32, 64, 128, 256])
def test_eig_hermitian(st, oracle, T):
    rng = np.random.default_rng(11)
    A = crand(rng, T, 3 * T).astype(np.complex128)
    A = (A @ A.conj().T).astype(np.complex64)
    ev, V = st.eig_hermitian(dev(A))
    ev, V = host(ev), host(V)
    w, Vo, _ = oracle.eig_jacobi(A.astype(np.complex128))
    assert np.allclose(ev, w, rtol=1e-10, atol=1e-10 * np.abs(w).max())
    # same phase convention -> same vectors (eigenvalues are distinct here)
    assert np.abs(V - Vo).max() < 1e-8
    A64 = A.astype(np.complex128)
    assert rel_l2(A64 @ V, V * ev[None, :]) < 1e-12


@pytest.mark.parametrize("T,pmin,pmax", [(8, 0, 7), (32, 4, 31), (16, 0, 0), (128, 100, 127), (128, 4, 60),
                                         (256, 0, 255)])
@pytest.mark.parametrize("img_type", [0, 1])
def test_pca_image(st, oracle, T, pmin, pmax, img_type):
    rng = np.random.default_rng(12)
    ring = crand(rng, T, 32, 32) * 50
    V = np.linalg.qr(rng.standard_normal((T, T)) + 1j * rng.standard_normal((T, T)))[0]
    got = host(st.pca_image(dev(ring), dev(V.astype(np.complex128)), img_type, pmin, pmax))
    C = oracle.pca_project(ring.astype(np.complex128), V)
    want = oracle.image_type(C, img_type, pmin, pmax) * 131071
    assert rel_l2(got, want) < 2e-6


@pytest.mark.parametrize("T,fs,fe", [(16, 0, 15), (32, 4, 31), (128, 4, 60), (256, 3, 200)])
@pytest.mark.parametrize("moment", [0, 1, 2])
def test_pca_moments(st, oracle, T, fs, fe, moment):
    """Moments of the PCA output (well-conditioned random unitary V): projection
    pinned against the oracle's matmul, moments against its fftfreq/Nyquist rule."""
    rng = np.random.default_rng(13)
    ring = crand(rng, T, 16, 32) * 50
    V = np.linalg.qr(rng.standard_normal((T, T)) + 1j * rng.standard_normal((T, T)))[0]
    fps = 1.5 * T
    got = host(st.pca_moments(dev(ring), dev(V.astype(np.complex128)), moment, fs, fe, fps))
    C = oracle.pca_project(ring.astype(np.complex128), V)
    w = oracle.moments(C, fs, fe, fps)
    # M1 mixes signs (negative frequencies): bound the error by M0's scale
    scale = np.abs(w[0]).max() * fps ** moment
    assert np.abs(got - w[moment]).max() < 2e-6 * scale


# (width, height): power-of-two sides run the four fused passes (square,
# wide, tall with 2048-long columns, the smallest), the others the step-by-step
# 2-D transforms
@pytest.mark.parametrize("W,H", [(256, 256), (512, 128), (64, 2048), (16, 16), (96, 80)])
def test_unwrap_2d(st, oracle, W, H):
    y, x = np.mgrid[0:H, 0:W]
    phi = 4.0 * np.sin(2 * np.pi * x / W) * np.cos(2 * np.pi * y / H) + 1.5 * np.cos(2 * np.pi * 2 * y / H)
    wrapped = np.angle(np.exp(1j * phi)).astype(np.float32)
    got = host(st.unwrap_2d(dev(wrapped)))
    want = oracle.unwrap2d(wrapped.astype(np.float64))
    assert rel_l2(got, want) < 1e-5
    r = host(st.rescale_unwrap(dev(got)))
    assert np.array_equal(r, oracle.rescale_unwrap(got))


@pytest.mark.parametrize("case", KATS["subsample"])
def test_subsample_kat(st, case):
    img = np.array(case["input"], np.float32).reshape(case["height"], case["width"])
    got = host(st.subsample(dev(img), case["step"]))
    assert got.ravel().tolist() == case["expected"]


@pytest.mark.parametrize("case", KATS["reduce"], ids=lambda c: f"{c['pattern']}-{c['width']}x{c['height']}")
def test_minmax_reference_patterns(st, case):
    n = case["width"] * case["height"]
    i = np.arange(n)
    if case["pattern"] == "mod10":
        x = (i % 10).astype(np.float32)
    elif case["pattern"] == "mod1000p1":
        x = ((i % 1000) + 1).astype(np.float32)
    else:
        x = -((i % 1000) + 1).astype(np.float32)
    r = host(st.minmax(dev(x)))
    assert r[0] == x.min() and r[1] == x.max()
    if case["op"] == "min":
        assert r[0] == case["expected"]
    if case["op"] == "max":
        assert r[1] == case["expected"]


@pytest.mark.parametrize("step", [1, 2, 8])
@pytest.mark.parametrize("lo,hi", [(0.5, 99.5), (0.02, 99.98), (0.0, 100.0), (50.0, 50.0)])
def test_percentile_exact(st, oracle, step, lo, hi):
    rng = np.random.default_rng(13)
    img = (rng.standard_normal((256, 128)) * 1000).astype(np.float32)
    img[3, 5] = -0.0
    img[7, 9] = 0.0
    got = host(st.percentile_range(dev(img), lo, hi, step, clamp=False))
    src = oracle.subsample(img, step) if step > 1 else img
    want = oracle.percentile(src, [lo, hi])
    assert np.array_equal(got, want)
    got_c = host(st.percentile_range(dev(img), lo, hi, step, clamp=True))
    assert np.array_equal(got_c, np.maximum(want, 1.0))


@pytest.mark.parametrize("H,W,step", [(256, 256, 2), (258, 256, 2), (64, 64, 1), (1024, 1024, 8)])
def test_percentile_one_workgroup_boundary(st, oracle, H, W, step):
    """Up to 16384 visited pixels the select runs in one workgroup on keys
    held in LDS (256^2 / 2: exactly 16384; 258 x 256 / 2: one row more, the
    multi-pass select); both equal the oracle's order statistics."""
    rng = np.random.default_rng(17)
    img = (rng.standard_normal((H, W)) * 1000).astype(np.float32)
    img[0, 0] = -0.0
    src = oracle.subsample(img, step) if step > 1 else img
    for lo, hi in [(0.5, 99.5), (0.0, 100.0), (50.0, 50.0)]:
        got = host(st.percentile_range(dev(img), lo, hi, step, clamp=False))
        assert np.array_equal(got, oracle.percentile(src, [lo, hi])), (lo, hi)


@pytest.mark.parametrize("W,H,scale", [(128, 256, 0.5), (256, 256, 1.0), (64, 32, 0.3), (512, 128, 0.0)])
def test_percentile_reticle_exact(st, oracle, W, H, scale):
    """Percentiles over the contrast reticle's centred disc equal the oracle's
    order statistics of the same pixels (scale 0: the centre pixel alone)."""
    rng = np.random.default_rng(21)
    img = (rng.standard_normal((H, W)) * 1000).astype(np.float32)
    disc = oracle.reticle_pixels(img, scale)
    for lo, hi in [(0.5, 99.5), (0.02, 99.98), (0.0, 100.0)]:
        got = host(st.percentile_range_reticle(dev(img), lo, hi, scale, clamp=False))
        assert np.array_equal(got, oracle.percentile(disc, [lo, hi])), (lo, hi)


def test_contrast_and_u16_exact(st, oracle):
    rng = np.random.default_rng(14)
    img = (rng.standard_normal(4099) * 30000 + 20000).astype(np.float32)
    img[:5] = [-1.0, 0.0, 65535.0, 65536.0, 1e9]
    t = dev(img)
    st.apply_contrast(t, 100.0, 50000.0)
    want = oracle.apply_contrast(img, 100.0, 50000.0)
    assert np.array_equal(host(t), want)
    u = st.u16(st.float_to_ushort(dev(want)))
    assert np.array_equal(u, oracle.float_to_u16(want))
    rng_dev = dev(np.array([100.0, 50000.0], np.float32))
    u2 = st.u16(st.contrast_to_ushort(dev(img), rng_dev))
    assert np.array_equal(u2, oracle.float_to_u16(want))
    u3 = st.u16(st.contrast_to_ushort(dev(img), rng_dev, invert=True))
    assert np.array_equal(u3, oracle.float_to_u16(oracle.apply_contrast(img, 50000.0, 100.0)))
