"""The next rows chained: a .holo recording read by the native file source
and read worker into the HBM batch input queue, handed to the pipe in place,
the u16 images recorded by the HBM record slots into a second .holo file --
which holds exactly the images the pipe makes from the same frames pushed
directly."""
import time

import numpy as np
import pytest
import torch

from oracle import holo_format as hf

pytestmark = pytest.mark.gpu


def test_recording_to_recording(tmp_path):
    from holovibes_amd import holo_file as hfm
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.input_queue import BatchInputQueue
    from holovibes_amd.pipe import Pipe, synth_frames
    from holovibes_amd.record import Recorder
    from holovibes_amd.sources import FrameReadWorker, HoloFileCamera
    s = settings_for("c1")
    B, H, W, n = s.batch_size, s.height, s.width, 5
    frames = synth_frames(synth_for("c1"), 0, n * B)
    src = str(tmp_path / "in.holo")
    with hfm.OutputHoloFile(src, hfm.FrameDescriptor(W, H, 2), n * B, hfm.RAW) as out:
        out.export_compute_settings(s, input_fps=1000, contiguous=n * B)
        out.write_header()
        out.write_frame(frames)
        out.write_footer()

    direct = Pipe(s)
    expect = []
    for b in range(n):
        assert direct.push(torch.from_numpy(frames[b * B:(b + 1) * B].view(np.int16)).cuda())
        expect.append(direct.output_u16())
    direct.close()

    # the file's own footer configures the pipe
    with hfm.InputHoloFile(src) as f:
        s2, _ = hfm.compute_settings_from_json(f.import_compute_settings(), f.get_frame_descriptor(),
                                               f.import_info(), base=s)
    q = BatchInputQueue(64 * B, B, W, H, 2)  # room for every batch: nothing is overwritten
    cam = HoloFileCamera(src, frames_per_call=7)  # a file buffer that does not divide the batch
    worker = FrameReadWorker(cam, q, fps_limit=20000.0)
    pipe = Pipe(s2)
    dst = str(tmp_path / "out.holo")
    rec = Recorder(dst, pipe, "hologram", n)
    t0 = time.time()
    done = 0
    while done < n:
        assert time.time() - t0 < 30, "no batch from the queue"
        b = q.front(timeout_ms=50)
        if b is None:
            continue
        made = pipe.push(b)
        q.pop()
        rec.after_push(b, made)
        done += 1
    worker.stop()
    assert rec.close(input_fps=1000) == n
    pipe.close()
    cam.shutdown()
    q.close()
    p = hf.parse(open(dst, "rb").read())
    assert (p["img_nb"], p["data_type"]) == (n, hf.PROCESSED)
    got = np.frombuffer(p["frames"], np.uint16).reshape(n, H, W)
    for k in range(n):
        assert np.array_equal(got[k], expect[k]), k
