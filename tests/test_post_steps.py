"""The post-processing restatement (oracle/post_steps.py) against what the
mathematics fixes: renormalised images have mean 2^c, a ramp of one image
averages to itself and constant images to their constant, means of
exactly representable values are exact, log10 of powers of ten is exact;
and the settings layout carries the new fields."""
import numpy as np

from oracle import post_steps as ps


def test_renormalize_mean_is_power_of_two():
    x = np.random.default_rng(0).random((64, 64)).astype(np.float32) * 37.0 + 1.0
    for c in (0, 5, 12):
        y = ps.renormalize(x, c)
        assert abs(float(y.mean(dtype=np.float64)) / 2 ** c - 1.0) < 1e-6
        # a scale of the input leaves the output unchanged up to rounding
        assert np.allclose(ps.renormalize(x * 4.0, c), y, rtol=1e-6)


def test_accumulator_means():
    acc = ps.Accumulator(3)
    imgs = [np.full((4, 4), v, np.float32) for v in (1.0, 2.0, 4.0, 8.0, 16.0)]
    got = [float(acc.push(im)[0, 0]) for im in imgs]
    # held images: 1, 2, 3, 3, 3 -> sums 1, 3, 7, 14, 28 (exact) over min(n, 3)
    third = [float(np.float32(v) / np.float32(3.0)) for v in (7.0, 14.0, 28.0)]
    assert got == [1.0, 1.5] + third
    one = ps.Accumulator(1)
    r = np.random.default_rng(1).random((8, 8)).astype(np.float32)
    assert np.array_equal(one.push(r), r)


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def test_log10_correctly_rounded():
    """The restatement is log10 rounded once to float: exact on powers of
    ten, within 2 ulp of numpy's float32 log10, monotone."""
    p10 = np.float32(10.0) ** np.arange(0, 8, dtype=np.float32)
    assert np.array_equal(ps.log10(p10), np.arange(0, 8, dtype=np.float32))
    x = np.sort(np.random.default_rng(2).random(4096).astype(np.float32) * 1e4 + 1e-3)
    assert int(ulps(ps.log10(x), np.log10(x)).max()) <= 2
    assert np.all(np.diff(ps.log10(x)) >= 0)


def test_settings_carry_post_steps():
    from holovibes_amd.configs import settings_for
    s = settings_for("c1", renorm_enabled=1, renorm_constant=7, img_accu_level=4, log_enabled=1)
    assert (s.renorm_enabled, s.renorm_constant, s.img_accu_level, s.log_enabled) == (1, 7, 4, 1)
    d = settings_for("c1")
    assert (d.renorm_enabled, d.renorm_constant, d.img_accu_level, d.log_enabled) == (0, 5, 1, 0)


def test_cuts_closed_forms_and_brute_force():
    T, H, W = 4, 6, 5
    # a constant field: every cut value is |c|, its square, or the argument
    c = np.full((T, H, W), 3 + 4j, np.complex64)
    xz, yz = ps.cuts(c, 1, 2, 3, 4, 0)
    assert np.all(xz == 5.0) and np.all(yz == 5.0) and xz.shape == (T, W) and yz.shape == (H, T)
    assert np.all(ps.cuts(c, 0, 0, 4, 5, 1)[0] == 25.0)
    pos = np.full((T, H, W), 2 + 0j, np.complex64)
    assert np.all(ps.cuts(pos, 0, 0, 0, 0, 2)[1] == np.float32(np.pi / 2))
    assert np.all(ps.cuts(c, 0, 0, 1, 1, 5)[0] == 0.0)  # moments: no value in the cuts
    # brute force on a small random window
    rng = np.random.default_rng(4)
    z = (rng.standard_normal((T, H, W)) + 1j * rng.standard_normal((T, H, W))).astype(np.complex64)
    xz, yz = ps.cuts(z, 1, 2, 3, 4, 0)
    for b in range(T):
        for x in range(W):
            acc = np.float32(0)
            for y in range(2, 5):
                acc = np.float32(acc + np.float32(abs(complex(z[b, y, x]))))
            assert abs(xz[b, x] - acc / np.float32(3)) <= 1e-6 * abs(acc)
        for y in range(H):
            acc = sum(abs(complex(z[b, y, x])) for x in range(1, 4)) / 3
            assert abs(yz[y, b] - acc) <= 1e-6 * acc


def test_registration_recovers_circular_shifts():
    """A smooth blob moved by a circular shift is moved back exactly, for
    shifts either side of the origin; the mask is a disc (swapped: its
    quadrants exchanged)."""
    H, W = 64, 96
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    img = (np.exp(-((x - 40) ** 2 + (y - 30) ** 2) / 60.0) + 0.5 * np.exp(-((x - 55) ** 2 + (y - 26) ** 2) / 20.0)
           ).astype(np.float32)
    mask = ps.registration_mask(W, H, 0.7, False)
    r = np.float32(32) * np.float32(0.7)
    assert mask[32, 48] == 1 and mask[0, 0] == 0 and int(mask.sum()) == int(
        (((x - 48) ** 2 + (y - 32) ** 2) <= r * r).sum())
    assert np.array_equal(ps.registration_mask(W, H, 0.7, True), np.fft.fftshift(mask))
    for dy, dx in [(3, -5), (-7, 2), (0, 0), (5, 9)]:
        moved = np.roll(img, (dy, dx), axis=(0, 1))
        out, (sx, sy) = ps.register(moved, img, mask)
        assert (sx, sy) == (-dx, -dy)
        assert np.array_equal(out, img)
