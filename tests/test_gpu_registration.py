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


def _run(s, batches):
    from holovibes_amd import _lib
    from holovibes_amd.pipe import Pipe
    p = Pipe(s)
    imgs, shifts = [], []
    for fr in batches:
        assert p.push(torch.from_numpy(fr.view(np.int16)).cuda())
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
