"""The reference's own CLI settings files drive the pipe: each file of
tests/golden/holo_settings/ whose computation is on the path, mapped onto
hv_compute_settings for 64x64 16-bit frames (the frame geometry comes from
the recording's header), runs on the GPU and on the oracle's pipe with the
same frames; images and u16 outputs agree to the pipe tests' bars (a file
that turns renormalisation on is checked against the oracle's image
renormalised by oracle/post_steps.py, its u16 not compared).  PCA
files select eigenvector 0 (the smallest eigenvalue, noise-dominated and
ill-conditioned on synthetic frames): their P range is moved to the last
component, the only change made to a file's settings."""
import glob
import json
import os

import numpy as np
import pytest
import torch

from oracle import post_steps as ps

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "holo_settings", "*.json")))


def _mapped(path):
    from holovibes_amd import holo_file as hfm
    cs = json.load(open(path))
    try:
        s, ignored = hfm.compute_settings_from_json(cs, hfm.FrameDescriptor(64, 64, 2), {"pixel_size": 20.0})
    except hfm.UnsupportedSettings:
        return None
    return s


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-5])
def test_settings_file_runs_like_oracle(oracle, path):
    from holovibes_amd.constants import HV_TIME_PCA
    from holovibes_amd.configs import synth_for
    from holovibes_amd.pipe import Pipe, synth_frames
    s = _mapped(path)
    if s is None:
        pytest.skip("computation outside the path (raw mode or SSA_STFT)")
    if s.time_transformation == HV_TIME_PCA:
        s.p_start, s.p_width = s.time_transformation_size - 1, 0
    B, S, T = s.batch_size, s.time_stride, s.time_transformation_size
    sp = synth_for("c1", width=64, height=64, period=T)
    for i in range(sp.n_discs):
        sp.disc_cx[i], sp.disc_cy[i], sp.disc_r[i] = 31.5, 31.5, 16.0
    n_push = (T + B - 1) // B * max(1, S // B) + 3 * max(1, S // B)
    print(f"{os.path.basename(path)}: B={B} S={S} T={T} space={s.space_transformation} "
          f"time={s.time_transformation} pushes={n_push}", flush=True)
    p = Pipe(s)
    o = oracle.OraclePipe(s, 64)
    tol = 1e-4 if s.time_transformation == HV_TIME_PCA else 1e-5
    made_any = 0
    for k in range(n_push):
        fr = synth_frames(sp, k * B, B)
        made = p.push(torch.from_numpy(fr.view(np.int16)).cuda())
        ref = o.push(fr)
        assert made == (ref is not None), k
        if made:
            img = p.image().cpu().numpy()
            u = p.output_u16()
            exp = ref["image"]
            if s.renorm_enabled:  # the oracle's pipe stops before the post-processing
                exp = ps.renormalize(exp, s.renorm_constant)
            e = float(np.linalg.norm(img - exp) / max(np.linalg.norm(exp), 1e-300))
            assert e < tol, (k, e)
            if not s.renorm_enabled:
                d = np.abs(u.astype(np.int64) - ref["u16"].astype(np.int64))
                assert d.max() <= 1 and (d == 0).mean() >= 0.95, (k, d.max(), (d == 0).mean())
            made_any += 1
    p.close()
    assert made_any >= 1
