"""Registration on the GPU: c1-geometry batches that are circular shifts of
the first batch give images that are (to FP32 rounding) circular shifts of
the first image; the pipe finds each shift (the same one the restatement in
oracle/post_steps.py finds on the unregistered images) and moves the image
back onto the first one."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import post_steps as ps

pytestmark = pytest.mark.gpu
SHIFTS = [(0, 0), (6, -9), (-4, 13), (11, 3)]  # (dy, dx) of each batch's frames


def _run(s, batches, graph=False):
    from holovibes_amd import _lib
    from holovibes_amd.pipe import Pipe
    p = Pipe(s, graph=graph)
    imgs, shifts = [], []
    # two input buffers reused in turn: from the second steady cycle on, graphs replay
    bufs = [torch.empty(batches[0].shape, dtype=torch.int16, device="cuda") for _ in range(2)]
    for k, fr in enumerate(batches):
        buf = bufs[k % 2]
        buf.copy_(torch.from_numpy(fr.view(np.int16)))
        torch.cuda.synchronize()
        assert p.push(buf)
        imgs.append(p.image().cpu().numpy())
        if s.registration_enabled:
            sx, sy = ctypes.c_int(), ctypes.c_int()
            _lib.call("hv_pipe_registration_shift", p.h, ctypes.byref(sx), ctypes.byref(sy))
            shifts.append((sx.value, sy.value))
    p.close()
    return imgs, shifts


@pytest.mark.parametrize("fft_shift", [0, 1])
def test_registration_moves_images_back(fft_shift):
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    base = settings_for("c1", fft_shift=fft_shift)
    reg = settings_for("c1", fft_shift=fft_shift, registration_enabled=1, registration_zone=0.7)
    first = synth_frames(synth_for("c1"), 0, base.batch_size)
    batches = [np.roll(first, (dy, dx), axis=(1, 2)).copy() for dy, dx in SHIFTS]
    plain, _ = _run(base, batches)
    got, shifts = _run(reg, batches)
    mask = ps.registration_mask(base.width, base.height, 0.7, fft_shift == 1)
    # cycle 0 takes the reference (the first image, unshifted)
    assert np.array_equal(got[0], plain[0])
    for k in range(1, len(SHIFTS)):
        dy, dx = SHIFTS[k]
        _, (osx, osy) = ps.register(plain[k], plain[0], mask)
        assert shifts[k] == (osx, osy) == (-dx, -dy), (k, shifts[k], (osx, osy))
        # the image moved back: the circular shift itself is exact
        assert np.array_equal(got[k], np.roll(plain[k], (-dy, -dx), axis=(0, 1)))
        err = float(np.abs(got[k] - plain[0]).max()) / float(np.abs(plain[0]).max())
        assert err < 1e-4, (k, err)


def test_registration_graph_replay_matches_eager():
    """With graphs on, the cycles before the reference is taken and the
    accumulation queue is full run eagerly, the later ones are replayed:
    images identical to the eager pipe's, cycle by cycle."""
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    s = settings_for("c1", registration_enabled=1, img_accu_level=2)
    first = synth_frames(synth_for("c1"), 0, s.batch_size)
    shifts = SHIFTS + [(2, 2), (-3, 5), (0, 0), (6, -9)]
    batches = [np.roll(first, (dy, dx), axis=(1, 2)).copy() for dy, dx in shifts]
    eager, se = _run(s, batches)
    replay, sr = _run(s, batches, graph=True)
    assert se == sr
    for a, b in zip(eager, replay):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("W,H", [(2048, 2048), (2048, 1024)])
def test_registration_large_frames(W, H):
    """At 2048 x 2048 the masked-mean and peak searches run past their
    partial-sum cap (1024 partials, more than 256 per thread in the final
    reduction); a non-square frame checks the row / column roles.  The
    shift equals the restatement's on the unregistered images and the image
    moves back exactly."""
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    over = dict(width=W, height=H, batch_size=2, time_stride=2, time_transformation_size=2, p_start=1, p_width=0)
    base = settings_for("c1", **over)
    reg = settings_for("c1", registration_enabled=1, registration_zone=0.7, **over)
    print(f"registration {W}x{H}", flush=True)
    first = synth_frames(synth_for("c1", width=W, height=H, period=2), 0, 2)
    shifts = [(0, 0), (37, -101), (-250, 9)]
    batches = [np.roll(first, (dy, dx), axis=(1, 2)).copy() for dy, dx in shifts]
    plain, _ = _run(base, batches)
    got, found = _run(reg, batches)
    mask = ps.registration_mask(W, H, 0.7, False)
    assert np.array_equal(got[0], plain[0])
    for k in range(1, len(shifts)):
        dy, dx = shifts[k]
        _, (osx, osy) = ps.register(plain[k], plain[0], mask)
        assert found[k] == (osx, osy) == (-dx, -dy), (k, found[k], (osx, osy))
        assert np.array_equal(got[k], np.roll(plain[k], (-dy, -dx), axis=(0, 1)))
