"""Optional post-processing in the pipe (renormalise, fft shift, image
accumulation, log10, in the reference's order) against the restatement in
oracle/post_steps.py applied to the images of the same pipe without them:
renormalisation and accumulation bit-exact (float32 in the reference's
operation order; the double sum rounds to the same float), log10 within
2 ulp of the correctly rounded value (the bound CUDA documents for the
reference's log10f), eager and graph replay alike."""
import numpy as np
import pytest
import torch

from oracle import post_steps as ps

pytestmark = pytest.mark.gpu
N_CYCLES = 6


def _images(settings, frames, graph=False):
    from holovibes_amd.pipe import Pipe
    p = Pipe(settings, graph=graph)
    out, u16 = [], []
    B = settings.batch_size
    for k in range(N_CYCLES):
        if p.push(torch.from_numpy(frames[k * B:(k + 1) * B].view(np.int16)).cuda()):
            out.append(p.image().cpu().numpy())
            u16.append(p.output_u16())
    p.close()
    return out, u16


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


@pytest.fixture(scope="module")
def base():
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    s = settings_for("c1")
    frames = synth_frames(synth_for("c1"), 0, N_CYCLES * s.batch_size)
    imgs, _ = _images(s, frames)
    assert len(imgs) == N_CYCLES
    return s, frames, imgs


@pytest.mark.parametrize("case", ["renorm", "accu3", "log", "all_shift", "accu3_graph"])
def test_post_steps_match_oracle(base, case):
    from holovibes_amd.configs import settings_for
    s0, frames, imgs = base
    over = {"renorm": dict(renorm_enabled=1, renorm_constant=5),
            "accu3": dict(img_accu_level=3),
            "log": dict(log_enabled=1),
            "all_shift": dict(renorm_enabled=1, renorm_constant=9, fft_shift=1, img_accu_level=4, log_enabled=1),
            "accu3_graph": dict(img_accu_level=3)}[case]
    print(f"post steps {case}", flush=True)
    s = settings_for("c1", **over)
    got, u16 = _images(s, frames, graph=case.endswith("graph"))
    acc = ps.Accumulator(s.img_accu_level) if s.img_accu_level > 1 else None
    for k, (g, a) in enumerate(zip(got, imgs)):
        e = a
        if s.renorm_enabled:
            e = ps.renormalize(e, s.renorm_constant)
        if s.fft_shift:
            e = ps.shift_corners(e)
        if acc:
            e = acc.push(e)
        if s.log_enabled:
            assert int(_ulps(g, ps.log10(e)).max()) <= 2, k
        else:
            assert np.array_equal(g, e), (case, k, float(np.abs(g - e).max()))
        # the contrast ran on the post-processed image: u16 spans the range
        assert u16[k].max() > 60000 and u16[k].min() < 10


def test_lanes_refused_with_accumulation():
    from holovibes_amd import _lib
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe
    s = settings_for("c1", img_accu_level=2)
    with pytest.raises(_lib.HoloError, match="one cycle at a time"):
        Pipe(s, lanes=2)
