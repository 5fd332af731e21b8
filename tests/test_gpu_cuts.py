"""XZ / YZ cuts on the GPU: the step (hv_stft_cuts) against the restatement
in oracle/post_steps.py on random windows for every image type (float sums
in the reference's order; the only differences are the float math
functions' last bits), and the pipe's cuts against the restatement applied
to a float64 STFT of the pipe's own ring."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import post_steps as ps

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("img_type", [0, 1, 2, 5])
def test_cuts_step_matches_oracle(img_type):
    from holovibes_amd import _lib
    T, H, W = 16, 64, 128
    rng = np.random.default_rng(img_type)
    z = (rng.standard_normal((T, H, W)) + 1j * rng.standard_normal((T, H, W))).astype(np.complex64)
    d = torch.from_numpy(z).cuda()
    xz = torch.empty((T, W), dtype=torch.float32, device="cuda")
    yz = torch.empty((H, T), dtype=torch.float32, device="cuda")
    xmin, ymin, xmax, ymax = 7, 20, 19, 41
    _lib.call("hv_stft_cuts", ctypes.c_void_p(xz.data_ptr()), ctypes.c_void_p(yz.data_ptr()),
              ctypes.c_void_p(d.data_ptr()), W, H, T, xmin, ymin, xmax, ymax, img_type,
              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    exz, eyz = ps.cuts(z, xmin, ymin, xmax, ymax, img_type)
    for got, exp in ((xz.cpu().numpy(), exz), (yz.cpu().numpy(), eyz)):
        assert np.allclose(got, exp, rtol=4e-6, atol=1e-6), float(np.abs(got - exp).max())
    with pytest.raises(_lib.HoloError, match="cuts"):
        _lib.call("hv_stft_cuts", ctypes.c_void_p(xz.data_ptr()), ctypes.c_void_p(yz.data_ptr()),
                  ctypes.c_void_p(d.data_ptr()), W, H, T, 0, 0, W, 5, img_type, None)


def test_pipe_cuts_match_stft_of_ring():
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import Pipe, synth_frames
    from holovibes_amd import _lib
    s = settings_for("c1", cuts_enabled=1, x_start=40, x_width=6, y_start=100, y_width=30)
    B, H, W, T = s.batch_size, s.height, s.width, s.time_transformation_size
    frames = synth_frames(synth_for("c1"), 0, 3 * B)
    p = Pipe(s)
    for b in range(3):
        made = p.push(torch.from_numpy(frames[b * B:(b + 1) * B].view(np.int16)).cuda())
        assert made
        p.sync()
        ring = p.ring()[0].cpu().numpy()
        xzp, yzp = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("hv_pipe_cuts", p.h, ctypes.byref(xzp), ctypes.byref(yzp))
        xz = torch.empty((T, W), dtype=torch.float32, device="cuda")
        yz = torch.empty((H, T), dtype=torch.float32, device="cuda")
        from holovibes_amd.pipe import copy_d2d
        copy_d2d(xz, xzp.value, xz.numel() * 4, p.stream)
        copy_d2d(yz, yzp.value, yz.numel() * 4, p.stream)
        p.sync()
        # the reference's STFT: unnormalised forward FFT along time, ring in slot order
        stft = np.fft.fft(ring.astype(np.complex128), axis=0).astype(np.complex64)
        exz, eyz = ps.cuts(stft, 40, 100, 46, 130, s.img_type)
        for got, exp in ((xz.cpu().numpy(), exz), (yz.cpu().numpy(), eyz)):
            err = float(np.abs(got - exp).max()) / float(np.abs(exp).max())
            assert err < 2e-5, (b, err)
    p.close()
