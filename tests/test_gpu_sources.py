"""Frame sources and the read worker on the GPU: the synthetic camera's
replayed ring (with its two-region wrap), a .holo recording looped over a
frame range with a file buffer that does not divide the batch, and a camera
written in Python through the C function table, each read by the native
worker thread into the HBM batch input queue; every batch taken from the
queue holds the frames the source produced, in order, and the frame-rate
limit holds."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _drain(q, fs):
    """Every whole batch left in the queue, as (frames, frame bytes)."""
    out = []
    while True:
        st = q.state()
        if st["size"] == 0:
            return out
        out.append(q.dequeue())


def _run(worker, want, timeout=20.0):
    t0 = time.time()
    while worker.frames() < want:
        assert time.time() - t0 < timeout, f"reader stalled at {worker.frames()} frames"
        time.sleep(0.005)
    return worker.stop()


def test_synthetic_camera_ring_and_rate():
    from holovibes_amd.configs import synth_for
    from holovibes_amd.input_queue import BatchInputQueue
    from holovibes_amd.pipe import synth_frames
    from holovibes_amd.sources import FrameReadWorker, SyntheticCamera
    p = synth_for("c1", width=64, height=64)
    ring, per_call, B = 20, 6, 8  # calls wrap the ring: 18..20 then 0..4
    ref = synth_frames(p, 0, ring).reshape(ring, -1).view(np.uint8)
    cam = SyntheticCamera(p, ring, per_call)
    q = BatchInputQueue(512, B, 64, 64, 2)
    fps = 3000.0
    t0 = time.perf_counter()
    w = FrameReadWorker(cam, q, fps_limit=fps)
    n = _run(w, 200)
    el = time.perf_counter() - t0
    assert n <= fps * el + per_call, (n, el)
    batches = _drain(q, 64 * 64 * 2)
    assert len(batches) == n // B
    for k, b in enumerate(batches):
        for j in range(B):
            assert np.array_equal(b[j], ref[(k * B + j) % ring]), (k, j)
    cam.shutdown()
    q.close()


def test_holo_file_source_loops_over_range(tmp_path):
    from holovibes_amd import holo_file as hfm
    from holovibes_amd.input_queue import BatchInputQueue
    from holovibes_amd.sources import FrameReadWorker, HoloFileCamera
    H = W = 32
    frames = np.random.default_rng(3).integers(0, 65536, (40, H, W), dtype=np.uint16)
    path = str(tmp_path / "src.holo")
    with hfm.OutputHoloFile(path, hfm.FrameDescriptor(W, H, 2), 40, hfm.RAW) as out:
        out.write_header()
        out.write_frame(frames)
    first, end, per_call, B = 5, 37, 7, 8
    cam = HoloFileCamera(path, per_call, first, end)
    q = BatchInputQueue(1024, B, W, H, 2)
    w = FrameReadWorker(cam, q, fps_limit=4000.0)
    n = _run(w, 100)
    batches = _drain(q, H * W * 2)
    order = [first + (i % (end - first)) for i in range(len(batches) * B)]
    # each call reads up to per_call frames and stops at `end`, so the loop order is the range repeated
    got = np.concatenate([b.view(np.uint16).reshape(B, H, W) for b in batches])
    assert np.array_equal(got, frames[order])
    assert len(batches) == n // B
    cam.shutdown()
    q.close()


def test_python_camera_plugin():
    from holovibes_amd.input_queue import BatchInputQueue
    from holovibes_amd.sources import FrameReadWorker, PyCamera

    class Counter(PyCamera):
        def __init__(self):
            super().__init__(16, 16, 2)
            self.k = 0

        def get_frames(self):
            a = np.empty((3, 16, 16), np.uint16)
            for j in range(3):
                a[j] = self.k + j
            self.k += 3
            return a

    cam = Counter()
    B = 4
    q = BatchInputQueue(256, B, 16, 16, 2)
    w = FrameReadWorker(cam, q, fps_limit=5000.0)
    n = _run(w, 60)
    batches = _drain(q, 16 * 16 * 2)
    got = np.concatenate([b.view(np.uint16).reshape(B, 16, 16)[:, 0, 0] for b in batches])
    assert np.array_equal(got, np.arange(len(got)))
    assert len(got) == n - n % B
    q.close()


def test_reader_stop_while_queue_full_and_front_held():
    """A full queue whose front batch the consumer holds blocks the read
    worker's next overwrite; stopping the worker must end that wait (the
    worker returns without an error) instead of deadlocking."""
    from holovibes_amd.configs import synth_for
    from holovibes_amd.input_queue import BatchInputQueue
    from holovibes_amd.sources import FrameReadWorker, SyntheticCamera
    p = synth_for("c1", width=32, height=32)
    B = 4
    cam = SyntheticCamera(p, 16, B)
    q = BatchInputQueue(2 * B, B, 32, 32, 2)  # two batch slots
    w = FrameReadWorker(cam, q)
    t0 = time.time()
    while q.state()["size"] < 2:
        assert time.time() - t0 < 20.0, "queue never filled"
        time.sleep(0.002)
    held = q.front(timeout_ms=1000)
    assert held is not None
    # the worker keeps overwriting the other slot until it reaches the held one
    time.sleep(0.2)
    n = w.stop()
    assert n >= 2 * B
    assert q.pop()
    cam.shutdown()
    q.close()
