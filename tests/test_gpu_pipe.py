"""GPU parity of the whole per-batch Pipe (hv_pipe_*) against the oracle's
Pipe (pipe.cc:295-455 restated), batch by batch, on seeded synthetic frames.

Bars: float image (before contrast) within relative L2 1e-5 of the FP64
oracle for FP32 paths (PCA split-BF16: 1e-4, the tolerance stated for the path); u16
output within 1 count on every pixel and identical on >= 95 % (truncation
to an integer flips a pixel whenever the ~1e-6 float error straddles an
integer); produced/discarded cycles identical.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    # in float64 / complex128: squared norms of PCA images (~1e20 per pixel) overflow float32
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def run_both(oracle, settings, frames_list, graph=False, fused=True, path=None):
    """fused: highest cycle path allowed (True = all); path: the path the
    pipe must have chosen (None: not checked)."""
    from holovibes_amd.pipe import Pipe
    p = Pipe(settings, graph=graph, fused=fused)
    if path is not None:
        assert p.path == path, (p.path, path)
    o = oracle.OraclePipe(settings, 64)
    results = []
    for fr in frames_list:
        # the kernels read batch_size frames of `depth` bytes per pixel: a mismatched array would be read past its end
        assert fr.dtype.itemsize == settings.depth and fr.size == settings.batch_size * settings.width * \
            settings.height, (fr.dtype, fr.shape, settings.depth)
        t = torch.from_numpy(fr.view(np.int16) if fr.dtype == np.uint16 else fr).cuda()
        torch.cuda.synchronize()
        made = p.push(t)
        ref = o.push(fr)
        assert made == (ref is not None)
        if made:
            img = p.image().cpu().numpy()
            u = p.output_u16()
            results.append((img, u, p.contrast_range(), ref))
    p.close()
    return results


def check(results, img_tol, u16_frac=0.95, u16_max=1):
    assert results, "no image produced"
    for img, u, rng, ref in results:
        e = rel_l2(img, ref["image"])
        assert e < img_tol, e
        d = np.abs(u.astype(np.int64) - ref["u16"].astype(np.int64))
        assert d.max() <= u16_max and (d == 0).mean() >= u16_frac, (d.max(), (d == 0).mean())


def frames(name, n_batches, **over):
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    s = settings_for(name, **{k: v for k, v in over.items() if not k.startswith("syn_")})
    syn = {"period": s.time_transformation_size}
    syn.update({k[4:]: v for k, v in over.items() if k.startswith("syn_")})
    sp = synth_for(name, width=s.width, height=s.height, **syn)
    cx, cy = (s.width - 1) / 2, (s.height - 1) / 2
    for i in range(sp.n_discs):
        if sp.disc_cx[i] >= s.width or sp.disc_cy[i] >= s.height:
            sp.disc_cx[i], sp.disc_cy[i] = cx, cy
        sp.disc_r[i] = min(sp.disc_r[i], s.width / 4)
    B = s.batch_size
    return s, [synth_frames(sp, k * B, B) for k in range(n_batches)]


@pytest.mark.parametrize("level", [2, 1, 0])
def test_pipe_c1_small(oracle, level):
    """B = S = T on each cycle path: bins-first (2), fused row-IFFT + STFT
    (1) and the three-pass + STFT sequence (0)."""
    s, fl = frames("c1", 3, width=64, height=64)
    check(run_both(oracle, s, fl, fused=level, path=level), 1e-5)


@pytest.mark.parametrize("W,H,T,p0,pw", [(64, 16, 2, 0, 1), (128, 64, 4, 1, 2), (256, 32, 8, 0, 7),
                                         (512, 16, 16, 2, 13), (1024, 16, 8, 3, 4), (64, 64, 64, 5, 40),
                                         (512, 32, 32, 0, 31), (1024, 32, 16, 15, 0)])
def test_pipe_bins_shapes(oracle, W, H, T, p0, pw):
    """Bins-first path over widths, window sizes and bin ranges (single bin,
    all bins, a range reaching T - 1)."""
    s, fl = frames("c1", 2, width=W, height=H, batch_size=T, time_stride=T, time_transformation_size=T,
                   p_start=p0, p_width=pw)
    check(run_both(oracle, s, fl, path=2), 1e-5)


@pytest.mark.parametrize("depth", [1, 2, 4])
@pytest.mark.parametrize("space", [1, 2])
def test_pipe_bins_depths(oracle, depth, space):
    """Bins-first path from u8 / u16 / f32 frames, Fresnel and angular spectrum."""
    syn = dict(syn_depth=1, syn_bits=8, syn_mean_counts=64.0) if depth == 1 else {}
    s, fl = frames("c1", 2, width=128, height=128, batch_size=16, time_stride=16, time_transformation_size=16,
                   depth=depth, space_transformation=space, z_distance=0.25, **syn)
    if depth == 4:
        fl = [f.astype(np.float32) * 0.5 for f in fl]
    check(run_both(oracle, s, fl, path=2), 1e-5)


@pytest.mark.parametrize("W,H,T", [(128, 64, 8), (256, 32, 32), (512, 16, 16), (1024, 16, 8)])
def test_pipe_fused_shapes(oracle, W, H, T):
    s, fl = frames("c1", 1, width=W, height=H, batch_size=T, time_stride=T, time_transformation_size=T,
                   p_start=1, p_width=T // 2 - 1)
    check(run_both(oracle, s, fl), 1e-5)


def test_pipe_ring_rotation_argument(oracle):
    """T = 2B: the window rotates in physical slot order; the argument image
    depends on that order (reference quirk 8.Q.6)."""
    from holovibes_amd.constants import HV_IMG_ARGUMENT
    s, fl = frames("c1", 5, width=64, height=64, batch_size=4, time_stride=4, time_transformation_size=8,
                   img_type=HV_IMG_ARGUMENT, p_start=1, p_width=0)
    res = run_both(oracle, s, fl)
    assert len(res) == 4
    check(res, 1e-4)


def test_pipe_stride_discard(oracle):
    s, fl = frames("c1", 6, width=32, height=32, batch_size=4, time_stride=8, time_transformation_size=4,
                   p_start=0, p_width=2)
    res = run_both(oracle, s, fl)
    assert len(res) == 3
    check(res, 1e-5)


@pytest.mark.parametrize("name,over,tol", [
    # B < T on the bins-first path: the raw frames of the window in their own ring
    ("c1", dict(width=64, height=64, batch_size=4, time_stride=4, time_transformation_size=16), 1e-5),
    # 3 does not divide 8: batches straddle the ring's end (two runs per write)
    ("c1", dict(width=64, height=64, batch_size=3, time_stride=3, time_transformation_size=8, p_start=1,
                p_width=4), 1e-5),
    # S = 2 B: every other batch dropped before the ring
    ("c1", dict(width=64, height=64, batch_size=4, time_stride=8, time_transformation_size=16), 1e-5),
    ("c1", dict(width=64, height=32, batch_size=8, time_stride=8, time_transformation_size=32, depth=1,
                space_transformation=1, z_distance=0.4, syn_depth=1, syn_bits=8, syn_mean_counts=64.0), 1e-5),
    ("c5_stft", dict(width=64, height=64, batch_size=16, time_stride=16, time_transformation_size=64, p_start=4,
                     p_width=40, camera_fps=64.0, syn_period=64), 1e-5),
    # split path: W = 2048 (time DFT, row, column and epilogue kernels), then column tiles (H = 2048)
    ("c1", dict(width=2048, height=16, batch_size=8, time_stride=8, time_transformation_size=32, p_start=2,
                p_width=9), 1e-5),
    ("c5_stft", dict(width=64, height=2048, batch_size=32, time_stride=32, time_transformation_size=128,
                     p_start=4, p_width=56, camera_fps=128.0, syn_period=128), 1e-5),
    # PCA of a sliding window: covariance, eigensolver and projection from the raw ring
    ("c2", dict(width=64, height=64, batch_size=8, time_stride=8, time_transformation_size=32), 1e-4),
    # K <= 4 on the split path: the direct few-bin transform inserts the batch into the raw ring as it reads
    ("c1", dict(width=2048, height=16, batch_size=8, time_stride=8, time_transformation_size=32, p_start=2,
                p_width=1), 1e-5),
    ("c1", dict(width=2048, height=16, batch_size=24, time_stride=24, time_transformation_size=64, p_start=0,
                p_width=3, as_f32=True), 1e-5),  # bin 0 (the window's first frame added back), float frames
    ("c1", dict(width=2048, height=16, batch_size=5, time_stride=5, time_transformation_size=16, p_start=3,
                p_width=0, depth=1, syn_depth=1, syn_bits=8, syn_mean_counts=64.0), 1e-5),
    # 1024-long columns with few planes (AS lens, Fresnel forward)
    ("c1", dict(width=64, height=1024, batch_size=8, time_stride=8, time_transformation_size=32, p_start=2,
                p_width=1), 1e-5),
    ("c1", dict(width=128, height=1024, batch_size=8, time_stride=8, time_transformation_size=32, p_start=1,
                p_width=0, space_transformation=1, z_distance=0.4), 1e-5),
])
def test_pipe_sliding_window_bins(oracle, name, over, tol):
    """Sliding windows (B < T) on the bins-first path, batch by batch
    against the oracle while the window fills and rotates."""
    over = dict(over)
    as_f32 = over.pop("as_f32", False)
    T, B = over["time_transformation_size"], over["batch_size"]
    # five images: the window fills after ceil(T / B) processed batches, one batch in S / B processed
    n = (-(-T // B) + 4) * (over["time_stride"] // B)
    if as_f32:
        over["depth"] = 4
    s, fl = frames(name, n, **over)
    if as_f32:
        fl = [f.astype(np.float32) * 0.5 for f in fl]
    res = run_both(oracle, s, fl, path=2)
    assert len(res) >= 5
    check(res, tol)


def test_pipe_oversized_batch(oracle):
    s, fl = frames("c1", 2, width=32, height=32, batch_size=8, time_stride=8, time_transformation_size=4,
                   p_start=1, p_width=2)
    check(run_both(oracle, s, fl), 1e-5)


@pytest.mark.parametrize("img_type", [0, 1, 2])
@pytest.mark.parametrize("space", [0, 1, 2])
def test_pipe_space_and_image_types(oracle, img_type, space):
    from holovibes_amd.constants import HV_CONTRAST_REFERENCE
    s, fl = frames("c1", 2, width=64, height=32, batch_size=8, time_stride=8, time_transformation_size=8,
                   space_transformation=space, img_type=img_type, p_start=1, p_width=3,
                   contrast_mode=HV_CONTRAST_REFERENCE, z_distance=0.25)
    check(run_both(oracle, s, fl), 1e-5 if img_type != 2 else 1e-4)


@pytest.mark.parametrize("level", [2, 0])
@pytest.mark.parametrize("precision,tol", [(0, 1e-5), (1, 1e-4), (2, 1e-4)])
def test_pipe_pca_fresnel(oracle, precision, tol, level):
    """c2 shape at 128 x 128: Fresnel + PCA (T=32), eigenvectors 4..31,
    squared modulus -- bins-first (2: covariance of the raw window, the 28
    selected components propagated) and the reference's step order (0)."""
    s, fl = frames("c2", 2, width=128, height=128, pca_precision=precision)
    check(run_both(oracle, s, fl, fused=level, path=level), tol, u16_frac=0.95)


@pytest.mark.parametrize("W,H,T,depth,space,p0,pw,precision", [
    (64, 64, 8, 2, 2, 2, 5, 0),       # angular spectrum, u16 (fused row kernel)
    (128, 64, 16, 1, 1, 4, 11, 1),    # Fresnel, u8, split BF16
    (128, 128, 64, 2, 2, 4, 59, 2),   # T = 64 (grouped covariance, fused rows), split FP16
    (2048, 16, 16, 2, 2, 4, 11, 1),   # W = 2048: projection, row and image as passes of their own
    (2048, 16, 32, 2, 2, 4, 27, 2),   # W = 2048, T = 32: the projection on split-FP16 MFMA
    (2048, 16, 64, 1, 1, 8, 55, 1),   # T = 64, u8, Fresnel: MFMA projection x chirp
    (2048, 32, 128, 2, 2, 4, 56, 2),  # T = 128, 57 components: four MFMA column tiles
])
def test_pipe_pca_bins_shapes(oracle, W, H, T, depth, space, p0, pw, precision):
    """The bins-first PCA path against the FP64 oracle pipe (which propagates
    every frame and takes the covariance of the propagated ring) on the
    synthetic interferograms, squared modulus of components p0 .. T - 1:
    integer depths, both propagators, the single-group and grouped covariance
    kernels, fused and split projection.  (The sum of squared moduli depends
    only on the components' span; the interferograms' time covariance has
    near-degenerate pairs whose individual eigenvectors are not determined.)"""
    syn = dict(syn_depth=2, syn_bits=12, syn_mean_counts=1024.0) if depth == 2 else {}  # c2's generator is u8
    s, fl = frames("c2", 2, width=W, height=H, depth=depth, batch_size=T, time_stride=T,
                   time_transformation_size=T, space_transformation=space, img_type=1, p_start=p0,
                   p_width=pw, pca_precision=precision, z_distance=0.25 if space == 2 else 0.4, syn_period=T, **syn)
    print(f"PCA bins {W}x{H} T={T} depth={depth} space={space} prec={precision}", flush=True)
    # T = 128: the float image stays within 5e-6 of FP64, but after the min/max contrast the
    # split-FP16 covariance's rounding moves up to 3 counts of the u16 image on a third of the
    # pixels (the same with the VALU projection)
    u16 = dict(u16_max=3, u16_frac=0.6) if T == 128 else {}
    check(run_both(oracle, s, fl, path=2), 1e-5 if precision == 0 else 1e-4, **u16)


@pytest.mark.parametrize("space", [1, 2])
@pytest.mark.parametrize("img_type,tol", [(0, 1e-5), (1, 1e-5), (2, 1e-4)])
def test_pipe_pca_bins_image_types(oracle, space, img_type, tol):
    """Modulus, squared modulus and argument of components 3 .. 8 of T = 16
    on the bins-first PCA path, on frames with a designed, well-separated
    time spectrum (the individual eigenvectors are then determined).  The
    argument image (mean phase of the projections) is within 1e-4 relative
    L2; its u16 image within 4 counts."""
    from holovibes_amd.configs import settings_for
    from holovibes_amd.constants import HV_DEPTH_32F
    T = 16
    s = settings_for("c2", width=64, height=32, depth=HV_DEPTH_32F, batch_size=T, time_stride=T,
                     time_transformation_size=T, space_transformation=space, img_type=img_type, p_start=3,
                     p_width=5, pca_precision=0, z_distance=0.25 if space == 2 else 0.4)
    check(run_both(oracle, s, [designed_frames(T, 64, 32, 5)], path=2), tol, u16_max=4 if img_type == 2 else 1)


def test_pipe_unwrap_percentile_subsample(oracle):
    """c3 shape at 128 x 128: argument + unwrap, percentile 0.5/99.5 on an 8x8 subsample."""
    s, fl = frames("c3", 3, width=128, height=128, time_transformation_size=16, batch_size=8, time_stride=8,
                   syn_period=16)
    res = run_both(oracle, s, fl)
    assert len(res) == 2
    # the unwrapped image is rescaled by its own min/max, so compare the u16 outputs
    check(res, 1e-3, u16_frac=0.95)


@pytest.mark.parametrize("path", [2, 0])
def test_pipe_minmax_partials_collapse(oracle, path):
    """1024 x 512 image: more than 256 contrast workgroups, so the per-block
    min/max partials are first reduced to one pair (k_mm_collapse); both the
    bins-first and the three-pass producers."""
    s, fl = frames("c1", 2, width=1024, height=512, batch_size=8, time_stride=8, time_transformation_size=8,
                   p_start=1, p_width=3)
    check(run_both(oracle, s, fl, fused=path, path=path), 1e-5)


def test_pipe_unwrap_rescale_collapse(oracle):
    """Unwrap at 512 x 256: the rescale's min/max has more than 256 consumer
    workgroups (partials collapsed first)."""
    s, fl = frames("c3", 2, width=512, height=256, time_transformation_size=8, batch_size=8, time_stride=8,
                   syn_period=8)
    check(run_both(oracle, s, fl), 1e-3, u16_frac=0.95)


@pytest.mark.parametrize("level", [2, 0])
@pytest.mark.parametrize("img_type", [5, 6, 7])
def test_pipe_moments(oracle, img_type, level):
    """Moments on the bins-first path (2: fused time DFT + row kernel, then the
    inverse row pass and the moment epilogue) and the three-pass path (0)."""
    s, fl = frames("c5_stft", 1, width=64, height=64, batch_size=32, time_stride=32, time_transformation_size=32,
                   img_type=img_type, p_start=4, p_width=20, camera_fps=32.0, syn_period=32)
    check(run_both(oracle, s, fl, fused=level, path=level), 1e-5)


@pytest.mark.parametrize("W,H,T,space,img_type,p0,pw", [
    (64, 64, 128, 2, 5, 4, 56),     # T = 128: the window's DFT in its own kernel, moment 0
    (64, 32, 256, 2, 6, 100, 60),   # T = 256, moment 1 over a range holding the Nyquist bin
    (2048, 16, 16, 2, 0, 2, 5),     # W = 2048: split row passes, modulus image
    (2048, 32, 8, 1, 1, 1, 3),      # Fresnel, W = 2048, squared modulus
    (128, 64, 128, 1, 5, 0, 127),   # Fresnel, T = 128, every bin (the Nyquist rule at the range edge)
    # 2048- and 4096-long columns: the planes column-tiled between the row passes (2 K <= T)
    (64, 2048, 128, 2, 5, 4, 56),   # four-column tiles, moment 0 over 57 bins
    (128, 4096, 128, 2, 0, 2, 9),   # two-column tiles, modulus image
    (64, 2048, 128, 2, 6, 40, 60),  # tiles, moment 1 over bins 40..100 (the Nyquist bin inside)
    (64, 2048, 128, 2, 7, 2, 30),   # tiles, moment 2
    (64, 2048, 64, 2, 2, 1, 20),    # tiles, argument image
    (32, 2048, 32, 2, 1, 0, 15),    # tiles, squared modulus, bins 0..15 (DC)
    (64, 2048, 64, 2, 6, 3, 40),    # 2 K > T: row-major planes, moment 1
    # the real window's half-length FFT (T >= 32) and the fused inverse row pass on row-major planes
    (2048, 16, 32, 2, 0, 2, 13),    # 512-pixel time-DFT tiles, modulus image
    (2048, 16, 64, 2, 5, 4, 40),    # 256-pixel tiles, moment 0
    (2048, 16, 64, 1, 1, 1, 30),    # Fresnel (x chirp), squared modulus
    (1080, 32, 32, 2, 0, 2, 13),    # frame not a multiple of 512 pixels: 64-pixel tiles
])
def test_pipe_bins_split(oracle, W, H, T, space, img_type, p0, pw):
    """The bins-first path as separate passes (frames wider than 1024 or
    windows longer than 64): time DFT of the raw window keeping bins
    [p_start, ...], row FFT, column pass, inverse row FFT, image or moment."""
    s, fl = frames("c5_stft", 1, width=W, height=H, batch_size=T, time_stride=T, time_transformation_size=T,
                   space_transformation=space, img_type=img_type, p_start=p0, p_width=pw, camera_fps=float(T),
                   z_distance=0.35, syn_period=T)
    # the argument's atan2 turns rounding at small |z| into larger angle steps (as for the other argument images)
    check(run_both(oracle, s, fl, path=2), 1e-5, u16_max=4 if img_type == 2 else 1)


def designed_frames(T, W, H, seed):
    """Float frames whose time covariance has well-separated eigenvalues:
    x_t = sum_k sqrt(l_k) u_k[t] b_k with u orthonormal over time, b
    orthonormal over pixels and l_k = 1 + k / T (relative gaps 1 / (2T)).
    The angular spectrum is unitary per frame, so the ring keeps them."""
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((T, T)))[0]
    B = np.linalg.qr(rng.standard_normal((W * H, T)))[0]
    lam = 1.0 + np.arange(T) / T
    X = (U * np.sqrt(lam)) @ B.T
    return (X * 1000).astype(np.float32).reshape(T, H, W)


@pytest.mark.parametrize("level", [2, 0])
@pytest.mark.parametrize("img_type", [5, 6, 7])
@pytest.mark.parametrize("T,precision,tol", [(32, 0, 1e-5), (128, 0, 1e-5), (128, 2, 1e-4)])
def test_pipe_pca_moments(oracle, img_type, T, precision, tol, level):
    """c5 PCA variant at 32 x 32: moments of the projections onto
    eigenvectors f_start..f_end (register kernel at T = 32, chunked projection
    over two component chunks at T = 128).  The synthetic interferograms'
    eigenvalues lie 1e-8 x lambda_max apart at T = 128, below FP32 covariance
    precision, so their eigenvectors are not determined; these frames have a
    designed, well-separated spectrum."""
    from holovibes_amd.configs import settings_for
    from holovibes_amd.constants import HV_DEPTH_32F
    s = settings_for("c5_pca", width=32, height=32, depth=HV_DEPTH_32F, batch_size=T, time_stride=T,
                     time_transformation_size=T, img_type=img_type, p_start=4, p_width=min(56, T - 5),
                     camera_fps=float(T), pca_precision=precision)
    check(run_both(oracle, s, [designed_frames(T, 32, 32, 17)], fused=level, path=level), tol)


@pytest.mark.parametrize("mode", [0, 1])
def test_pipe_contrast_reticle(oracle, mode):
    """Autocontrast over the reticle disc (ContrastReticleDisplayEnabled,
    rendering.cc:310-327): the reference policy (first window, then frozen)
    and percentiles on every image."""
    s, fl = frames("c1", 3, width=128, height=64, contrast_mode=mode, contrast_reticle_enabled=1,
                   contrast_reticle_scale=0.6, contrast_lower=0.5, contrast_upper=99.5)
    res = run_both(oracle, s, fl)
    check(res, 1e-5)
    for _, _, rng, ref in res:
        assert np.allclose(rng, ref["range"], rtol=1e-5)


def test_pipe_fft_shift_and_invert(oracle):
    s, fl = frames("c1", 1, width=64, height=64, fft_shift=1, contrast_invert=1)
    check(run_both(oracle, s, fl), 1e-5)


def test_pipe_graph_replay_equals_eager(oracle):
    """hipGraph replays give bit-identical outputs to the eager launches."""
    from holovibes_amd.pipe import Pipe
    s, fl = frames("c1", 6, width=64, height=64, batch_size=4, time_stride=4, time_transformation_size=8)
    eager, graph = Pipe(s), Pipe(s, graph=True)
    for fr in fl:
        t = torch.from_numpy(fr.view(np.int16)).cuda()
        torch.cuda.synchronize()
        a, b = eager.push(t), graph.push(t)
        assert a == b
        if a:
            assert np.array_equal(eager.output_u16(), graph.output_u16())
            assert torch.equal(eager.image(), graph.image())


@pytest.mark.parametrize("level", [2, 1, 0])
def test_pipe_metric_full_size(oracle, level):
    """The benchmark workload at full size (512 x 512, T = B = S = 32) on each cycle path."""
    s, fl = frames("metric", 1)
    check(run_both(oracle, s, fl, graph=True, fused=level, path=level), 1e-5)


@pytest.mark.parametrize("lanes", [2, 3])
@pytest.mark.parametrize("contrast_mode", [0, 2])
@pytest.mark.parametrize("kind", ["stft", "pca"])
def test_pipe_lanes_match_one_lane(lanes, contrast_mode, kind):
    """Cycles side by side on several lanes give the one-lane outputs,
    batch by batch and after a burst of pushes without synchronisation
    (min/max contrast per cycle; autocontrast frozen after the first); STFT
    bins and PCA components (each lane with its own covariance, eigensolver
    and projection buffers)."""
    from holovibes_amd.pipe import Pipe
    if kind == "stft":
        s, fl = frames("c1", 6, width=128, height=128, batch_size=16, time_stride=16, time_transformation_size=16,
                       contrast_mode=contrast_mode, contrast_lower=0.5, contrast_upper=99.5)
    else:
        s, fl = frames("c2", 6, width=128, height=128, contrast_mode=contrast_mode, contrast_lower=0.5,
                       contrast_upper=99.5)
    dev = [torch.from_numpy(f.view(np.int16) if f.dtype == np.uint16 else f).cuda() for f in fl]
    torch.cuda.synchronize()
    one, many = Pipe(s, graph=True), Pipe(s, graph=True, lanes=lanes)
    assert one.path == Pipe.BINS and many.path == Pipe.BINS
    for d in dev:
        assert one.push(d) and many.push(d)
        assert np.array_equal(one.output_u16(), many.output_u16())
        assert torch.equal(one.image(), many.image())
    for d in dev + dev[::-1]:
        one.push(d)
        many.push(d)
    assert np.array_equal(one.output_u16(), many.output_u16())
    assert np.array_equal(one.contrast_range(), many.contrast_range())
    one.close()
    many.close()


@pytest.mark.parametrize("case", ["stft", "row_stft", "pca", "pca_moments", "unwrap", "shared_stft",
                                  "shared_unwrap", "shared_pca"])
@pytest.mark.parametrize("lanes", [2, 3])
def test_pipe_lanes_other_paths_match_one_lane(case, lanes):
    """The other cycle paths on several lanes give the one-lane outputs: with
    B = S = T every batch replaces the window and each lane has its own ring
    and PCA buffers; with B < T the lanes share the ring."""
    from holovibes_amd.pipe import Pipe
    cfg = {"stft": ("c1", 0, dict(width=128, height=128, batch_size=16, time_stride=16, time_transformation_size=16)),
           "row_stft": ("c1", 1, dict(width=128, height=128, batch_size=16, time_stride=16,
                                      time_transformation_size=16)),
           "pca": ("c2", 0, dict(width=128, height=128)),
           "pca_moments": ("c5_pca", 0, dict(width=64, height=64, batch_size=32, time_stride=32,
                                             time_transformation_size=32, p_start=4, p_width=20, camera_fps=32.0,
                                             syn_period=32)),
           "unwrap": ("c3", 0, dict(width=128, height=128, batch_size=16, time_stride=16,
                                    time_transformation_size=16, syn_period=16)),
           # B < T: the lanes share the ring (each cycle waits for the previous cycle's time transform)
           "shared_stft": ("c1", 0, dict(width=128, height=64, batch_size=4, time_stride=4,
                                         time_transformation_size=16)),
           "shared_unwrap": ("c3", 0, dict(width=128, height=128, batch_size=8, time_stride=8,
                                           time_transformation_size=16, syn_period=16)),
           "shared_pca": ("c2", 0, dict(width=128, height=128, batch_size=8, time_stride=8,
                                        time_transformation_size=32))}[case]
    name, level, over = cfg
    s, fl = frames(name, 8, **over)
    conv = (lambda f: torch.from_numpy(f.view(np.int16) if f.dtype == np.uint16 else f).cuda())
    dev = [conv(f) for f in fl]
    torch.cuda.synchronize()
    one, many = Pipe(s, graph=True, fused=level), Pipe(s, graph=True, fused=level, lanes=lanes)
    assert one.path == many.path and one.path != Pipe.BINS
    for d in dev:
        a, b = one.push(d), many.push(d)
        assert a == b
        if a:
            assert np.array_equal(one.output_u16(), many.output_u16())
            assert torch.equal(one.image(), many.image())
    for d in dev + dev[::-1]:
        one.push(d)
        many.push(d)
    assert np.array_equal(one.output_u16(), many.output_u16())
    assert np.array_equal(one.contrast_range(), many.contrast_range())
    one.close()
    many.close()


@pytest.mark.parametrize("case", ["stft", "unwrap", "pca"])
@pytest.mark.parametrize("lanes", [2, 3])
def test_pipe_sliding_bins_lanes_match_one_lane(case, lanes):
    """Sliding windows on the bins-first path share the raw ring across the
    lanes (each cycle's write waits until the previous cycle's time DFT has
    read it): the lanes' outputs are bit-identical to one lane's."""
    from holovibes_amd.pipe import Pipe
    name, over = {"stft": ("c1", dict(width=128, height=64, batch_size=4, time_stride=4,
                                      time_transformation_size=16)),
                  "unwrap": ("c3", dict(width=128, height=128, batch_size=8, time_stride=8,
                                        time_transformation_size=16, syn_period=16)),
                  "pca": ("c2", dict(width=128, height=128, batch_size=8, time_stride=8,
                                     time_transformation_size=32))}[case]
    s, fl = frames(name, 8, **over)
    conv = (lambda f: torch.from_numpy(f.view(np.int16) if f.dtype == np.uint16 else f).cuda())
    dev = [conv(f) for f in fl]
    torch.cuda.synchronize()
    one, many = Pipe(s, graph=True), Pipe(s, graph=True, lanes=lanes)
    assert one.path == many.path == Pipe.BINS
    for d in dev:
        a, b = one.push(d), many.push(d)
        assert a == b
        if a:
            assert np.array_equal(one.output_u16(), many.output_u16())
            assert torch.equal(one.image(), many.image())
    for d in dev + dev[::-1]:
        one.push(d)
        many.push(d)
    assert np.array_equal(one.output_u16(), many.output_u16())
    assert torch.equal(one.image(), many.image())
    one.close()
    many.close()


def test_pipe_sliding_bins_path_switch(oracle):
    """Leaving the sliding-window bins-first path mid-stream (hv_pipe_set_fused
    0): the three-pass ring is rebuilt from the raw window and the next
    cycles match the oracle; coming back, the window refills first."""
    from holovibes_amd.pipe import Pipe
    from holovibes_amd import _lib
    s, fl = frames("c1", 10, width=64, height=64, batch_size=4, time_stride=4, time_transformation_size=16)
    p = Pipe(s)
    o = oracle.OraclePipe(s, 64)
    assert p.path == Pipe.BINS
    dev = [torch.from_numpy(f.view(np.int16)).cuda() for f in fl]
    torch.cuda.synchronize()
    for k in range(6):
        assert p.push(dev[k]) == (o.push(fl[k]) is not None)
    _lib.call("hv_pipe_set_fused", p.h, 0)
    assert p.path == Pipe.THREE_PASS
    for k in range(6, 8):
        ref = o.push(fl[k])
        assert p.push(dev[k]) and ref is not None
        assert rel_l2(p.image().cpu().numpy(), ref["image"]) < 1e-5
    _lib.call("hv_pipe_set_fused", p.h, 2)
    assert p.path == Pipe.BINS
    # the window refills: T / B - 1 pushes without an image
    assert not p.push(dev[8]) and not p.push(dev[9])
    p.close()


@pytest.mark.parametrize("lanes", [2, 3])
def test_pipe_bins_column_tiles_lanes(lanes):
    """Column-tiled bins (each extra lane holds its planes and tiles): the
    lanes' outputs are bit-identical to one lane's."""
    from holovibes_amd.pipe import Pipe
    s, fl = frames("c5_stft", 4, width=64, height=2048, batch_size=128, time_stride=128, time_transformation_size=128,
                   p_start=4, p_width=56, camera_fps=128.0, syn_period=128)
    dev = [torch.from_numpy(f.view(np.int16)).cuda() for f in fl]
    torch.cuda.synchronize()
    one, many = Pipe(s, fused=2), Pipe(s, fused=2, lanes=lanes)
    assert one.path == many.path == Pipe.BINS
    for d in dev + dev[::-1]:
        assert one.push(d) and many.push(d)
        assert np.array_equal(one.output_u16(), many.output_u16())
        assert torch.equal(one.image(), many.image())
    one.close()
    many.close()


@pytest.mark.parametrize("lanes", [1, 3])
def test_pipe_push_batches_matches_push(lanes):
    """hv_pipe_push_batches runs the cycles of per-batch pushes: the same
    number of images and the same last image (ring still filling at first)."""
    from holovibes_amd.pipe import Pipe
    s, fl = frames("c1", 7, width=64, height=64, batch_size=8, time_stride=8, time_transformation_size=16)
    dev = [torch.from_numpy(f.view(np.int16)).cuda() for f in fl]
    torch.cuda.synchronize()
    one, many = Pipe(s, lanes=lanes), Pipe(s, lanes=lanes)
    made = sum(bool(one.push(d)) for d in dev)
    assert made == len(dev) - 1
    assert many.push_many(Pipe.batch_pointers(dev)) == made
    assert np.array_equal(one.output_u16(), many.output_u16())
    assert torch.equal(one.image(), many.image())
    one.close()
    many.close()


def test_pipe_lanes_need_own_streams():
    from holovibes_amd import _lib
    from holovibes_amd.pipe import Pipe
    s, _ = frames("c1", 1, width=64, height=64, batch_size=4, time_stride=4, time_transformation_size=8)
    with torch.cuda.stream(torch.cuda.Stream()):
        p = Pipe(s)  # on the caller's (non-default) stream
    assert _lib.status("hv_pipe_set_lanes", p.h, 2) == _lib.HV_ERR_UNSUPPORTED
    p.close()


def test_server_alone_vs_mix():
    """Continuous batching: a source's outputs are identical alone and when
    its batches interleave with other sources' on the device."""
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import Server, synth_frames
    sa = settings_for("c1", width=128, height=128, batch_size=8, time_stride=8, time_transformation_size=8)
    sb = settings_for("metric", width=256, height=256, batch_size=16, time_stride=16, time_transformation_size=16)
    fa = [torch.from_numpy(synth_frames(synth_for("c1", width=128, height=128, period=8), 8 * k, 8).view(np.int16))
          .cuda() for k in range(4)]
    fb = [torch.from_numpy(synth_frames(synth_for("metric", width=256, height=256, period=16), 16 * k, 16)
                           .view(np.int16)).cuda() for k in range(4)]
    torch.cuda.synchronize()
    alone = Server(1 << 30)
    a = alone.add_source(sa)
    outs_alone = []
    for f in fa:
        alone.submit(a, f.data_ptr(), 0)
        alone.drain()
        outs_alone.append(alone.output_u16(a, 128, 128))
    alone.close()
    mix = Server(1 << 30)
    a2, b2 = mix.add_source(sa), mix.add_source(sb)
    outs_mix = []
    for fa_k, fb_k in zip(fa, fb):
        mix.submit(b2, fb_k.data_ptr(), 0)
        mix.submit(a2, fa_k.data_ptr(), 0)
        mix.drain()
        outs_mix.append(mix.output_u16(a2, 128, 128))
    lat = mix.poll()
    assert len(lat) == 8 and all(l > 0 for _, l in lat)
    mix.close()
    for x, y in zip(outs_alone, outs_mix):
        assert np.array_equal(x, y)


def test_server_budget_admission():
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe, Server
    s = settings_for("metric")
    need = Pipe.memory_bytes(s)
    srv = Server(int(need * 2.5))
    assert srv.add_source(s) is not None and srv.add_source(s) is not None
    assert srv.add_source(s) is None  # third does not fit
    srv.remove_source(0)
    assert srv.add_source(s) is not None  # admitted once memory is released
    assert srv.memory_bytes() == 2 * need
    srv.close()
