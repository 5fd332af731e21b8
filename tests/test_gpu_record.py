"""The record path on the GPU: moments (STFT and PCA), the u16 hologram and
raw batches written from HBM into .holo files, read back by the format
restatement: the moment frames equal the pipe's own moment images (each
parity-tested against the oracle), frame offset and skip select the
reference's frames, and the header is corrected to the frames recorded."""
import numpy as np
import pytest
import torch

from oracle import holo_format as hf

pytestmark = pytest.mark.gpu


def _frames(name, n, **over):
    from holovibes_amd.configs import synth_for
    from holovibes_amd.pipe import synth_frames
    return synth_frames(synth_for(name, **over), 0, n)


def _dev(fr):
    return torch.from_numpy(fr.view(np.int16) if fr.dtype == np.uint16 else fr).cuda()


def _moment_images(s, frames, cycles):
    """Each moment as the pipe's own image type, cycle by cycle, on the
    three-pass path the record runs (the bins-first moments round
    differently; both are parity-tested against the oracle)."""
    from holovibes_amd.configs import settings_for  # noqa: F401
    from holovibes_amd.constants import HV_IMG_MOMENTS_0
    from holovibes_amd.pipe import Pipe
    out = [[] for _ in range(cycles)]
    B = s.batch_size
    for k in range(3):
        s2 = type(s).from_buffer_copy(s)
        s2.img_type = HV_IMG_MOMENTS_0 + k
        p = Pipe(s2, fused=0)
        c = 0
        for b in range(cycles):
            if p.push(_dev(frames[b * B:(b + 1) * B])):
                out[c].append(p.image().cpu().numpy())
                c += 1
        p.close()
    return out


@pytest.mark.parametrize("name,over", [("c1", dict(camera_fps=64.0, p_start=2, p_width=5)),
                                       ("c2", dict(camera_fps=100.0, p_start=4, p_width=20))])
def test_moments_record(tmp_path, name, over):
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe
    from holovibes_amd.record import Recorder
    print(f"moments record {name}", flush=True)
    s = settings_for(name, **over)
    cycles = 3
    frames = _frames(name, cycles * s.batch_size)
    expect = _moment_images(s, frames, cycles)
    pipe = Pipe(s)
    path = str(tmp_path / "m.holo")
    rec = Recorder(path, pipe, "moments", 3 * cycles)
    images = []
    for b in range(cycles):
        batch = _dev(frames[b * s.batch_size:(b + 1) * s.batch_size])
        made = pipe.push(batch)
        rec.after_push(batch, made)
        if made:
            images.append(pipe.output_u16())
    assert rec.close(input_fps=1000) == 3 * cycles
    # the displayed image is unchanged by the record
    pipe.close()
    p = hf.parse(open(path, "rb").read())
    assert (p["img_nb"], p["depth"], p["data_type"]) == (3 * cycles, 4, hf.MOMENTS)
    got = np.frombuffer(p["frames"], np.float32).reshape(cycles, 3, s.height, s.width)
    for c in range(cycles):
        for k in range(3):
            assert np.array_equal(got[c, k], expect[c][k]), (c, k)
    assert len(images) == cycles


def test_hologram_record_offset_and_skip(tmp_path):
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe
    from holovibes_amd.record import Recorder
    s = settings_for("c1")
    n = 7
    frames = _frames("c1", n * s.batch_size)
    pipe = Pipe(s)
    path = str(tmp_path / "h.holo")
    rec = Recorder(path, pipe, "hologram", 10, frame_offset=1, frame_skip=1)
    outs = []
    for b in range(n):
        batch = _dev(frames[b * s.batch_size:(b + 1) * s.batch_size])
        made = pipe.push(batch)
        rec.after_push(batch, made)
        if made:
            outs.append(pipe.output_u16())
    assert rec.close(input_fps=100) == 3  # outputs 1, 3, 5 of 7
    pipe.close()
    p = hf.parse(open(path, "rb").read())
    assert (p["img_nb"], p["depth"], p["data_type"]) == (3, 2, hf.PROCESSED)
    got = np.frombuffer(p["frames"], np.uint16).reshape(3, s.height, s.width)
    for j, k in enumerate((1, 3, 5)):
        assert np.array_equal(got[j], outs[k])
    import json
    meta = json.loads(p["footer"])
    assert meta["info"]["contiguous"] == 3 and meta["info"]["input_fps"] == 50


def test_raw_record_and_frame_count_correction(tmp_path):
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe
    from holovibes_amd.record import Recorder
    s = settings_for("c1")
    frames = _frames("c1", 3 * s.batch_size)
    pipe = Pipe(s)
    path = str(tmp_path / "r.holo")
    rec = Recorder(path, pipe, "raw", 100)
    for b in range(3):
        batch = _dev(frames[b * s.batch_size:(b + 1) * s.batch_size])
        rec.after_push(batch, pipe.push(batch))
    assert rec.close(input_fps=1000) == 48
    pipe.close()
    p = hf.parse(open(path, "rb").read())
    assert (p["img_nb"], p["depth"], p["data_type"]) == (48, 2, hf.RAW)
    assert p["frames"] == frames.tobytes()


def test_moments_record_close_restores_graph_pipe(tmp_path):
    """Closing a moments record on a graph pipe drops the graphs captured
    with the record (they write into its released frames) and returns to
    the path the record displaced; later cycles equal an eager pipe's."""
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe
    from holovibes_amd.record import Recorder
    s = settings_for("c1", width=128, height=128)
    n = 6
    frames = _frames("c1", n * s.batch_size, width=128, height=128)
    batches = [_dev(frames[b * s.batch_size:(b + 1) * s.batch_size]) for b in range(n)]
    torch.cuda.synchronize()
    graph, eager = Pipe(s, graph=True), Pipe(s)
    assert graph.path == Pipe.BINS
    rec = Recorder(str(tmp_path / "m.holo"), graph, "moments", 3 * 3)
    assert graph.path == Pipe.THREE_PASS
    for b in batches[:3]:
        made = graph.push(b)
        rec.after_push(b, made)
        eager.push(b)
    assert rec.close(input_fps=100) == 9
    assert graph.path == Pipe.BINS
    for b in batches[3:]:
        assert graph.push(b) and eager.push(b)
        assert np.array_equal(graph.output_u16(), eager.output_u16())
        assert torch.equal(graph.image(), eager.image())
    graph.close()
    eager.close()


def test_moments_record_sliding_window_bins(tmp_path):
    """A sliding window (B < T) on the bins-first path holds the raw frames
    only: a moments record started mid-stream rebuilds the three-pass ring
    from them (the recorded moments equal a three-pass pipe's), the raw ring
    follows the window meanwhile, and after the record the bins-first
    cycles equal those of a pipe that never left the path."""
    from holovibes_amd.configs import settings_for
    from holovibes_amd.constants import HV_IMG_MOMENTS_0
    from holovibes_amd.pipe import Pipe
    from holovibes_amd.record import Recorder
    s = settings_for("c1", width=128, height=128, batch_size=8, time_stride=8, time_transformation_size=16,
                     camera_fps=64.0, p_start=2, p_width=5)
    B, n = s.batch_size, 8
    frames = _frames("c1", n * B, width=128, height=128)
    batches = [_dev(frames[b * B:(b + 1) * B]) for b in range(n)]
    torch.cuda.synchronize()
    # the moments of cycles 1..3 (batches 2..4) from three-pass pipes
    expect = []
    for k in range(3):
        s2 = type(s).from_buffer_copy(s)
        s2.img_type = HV_IMG_MOMENTS_0 + k
        p = Pipe(s2, fused=0)
        imgs = [p.image().cpu().numpy() for b in batches[:5] if p.push(b)]
        p.close()
        assert len(imgs) == 4
        expect.append(imgs[1:])
    graph, eager = Pipe(s, graph=True), Pipe(s)
    assert graph.path == Pipe.BINS
    for b in batches[:2]:
        graph.push(b)
        eager.push(b)
    rec = Recorder(str(tmp_path / "m.holo"), graph, "moments", 3 * 3)
    assert graph.path == Pipe.THREE_PASS
    for b in batches[2:5]:
        made = graph.push(b)
        assert made
        rec.after_push(b, made)
        eager.push(b)
    assert rec.close(input_fps=100) == 9
    assert graph.path == Pipe.BINS
    for b in batches[5:]:
        assert graph.push(b) and eager.push(b)
        assert np.array_equal(graph.output_u16(), eager.output_u16())
        assert torch.equal(graph.image(), eager.image())
    graph.close()
    eager.close()
    p = hf.parse(open(str(tmp_path / "m.holo"), "rb").read())
    got = np.frombuffer(p["frames"], np.float32).reshape(3, 3, s.height, s.width)
    for c in range(3):
        for k in range(3):
            assert np.array_equal(got[c, k], expect[k][c]), (c, k)


def test_recorder_refuses_lanes(tmp_path):
    """The record copies a cycle's output on the pipe's first stream, so it
    needs the cycles in order on one lane."""
    from holovibes_amd.configs import settings_for
    from holovibes_amd.pipe import Pipe
    from holovibes_amd.record import Recorder
    s = settings_for("c1", width=64, height=64)
    p = Pipe(s, lanes=2)
    for mode in ("raw", "hologram", "moments"):
        with pytest.raises(ValueError, match="one lane"):
            Recorder(str(tmp_path / f"{mode}.holo"), p, mode, 4)
    p.close()
