""".holo recordings on the GPU: frames streamed from a file into HBM
(hv_holo_load_frames) and from HBM into a file (hv_holo_store_frames) are
the bytes the format restatement (oracle/holo_format.py) reads and writes,
across the 32 MiB staging chunks; and a recording replayed through the pipe
with the settings of its own footer gives the output the same frames give
when pushed directly, recorded back into a .holo file bit for bit."""
import numpy as np
import pytest
import torch

from oracle import holo_format as hf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def holo():
    from holovibes_amd import holo_file
    holo_file._lib.load()
    return holo_file


def _frames(n, h, w, depth, seed):
    rng = np.random.default_rng(seed)
    if depth == 4:
        return rng.standard_normal((n, h, w)).astype(np.float32)
    dt = np.uint8 if depth == 1 else np.uint16
    return rng.integers(0, np.iinfo(dt).max + 1, (n, h, w), dtype=dt)


# (frames, height, width, depth, first, count): one chunk; several chunks with
# a tail; a 32 MiB chunk boundary inside a frame; odd sizes with a count past
# the end; 8-bit frames over two chunks
LOAD_CASES = [(3, 64, 64, 2, 0, 3), (40, 1024, 1024, 2, 2, 37), (20, 1000, 1000, 2, 3, 17),
              (9, 33, 17, 4, 4, 10), (130, 512, 512, 1, 1, 129)]


@pytest.mark.parametrize("n,h,w,depth,first,count", LOAD_CASES)
def test_load_frames_match_file(holo, tmp_path, n, h, w, depth, first, count):
    print(f"load n={n} {h}x{w} depth={depth} first={first} count={count}", flush=True)
    frames = _frames(n, h, w, depth, seed=n + h)
    data = hf.holo_bytes(frames, depth, hf.RAW, footer={"info": {"input_fps": 1}})
    path = str(tmp_path / "in.holo")
    open(path, "wb").write(data)
    expect = hf.parse(data)["frames"][first * h * w * depth:]
    with holo.InputHoloFile(path) as f:
        dev = torch.full((count, h * w * depth), 0xAB, dtype=torch.uint8, device="cuda")
        f.set_pos_to_frame(first)
        got = f.read_frames_to_device(dev, count, torch.cuda.current_stream())
        assert got == min(count, n - first)
        host = dev[:got].cpu().numpy().tobytes()
        assert host == expect[: got * h * w * depth]
        # frames past the end of the recording are left untouched
        assert bool((dev[got:] == 0xAB).all())
        t = f.load_to_device()
        assert t.shape == (n, h * w * depth) and t.cpu().numpy().tobytes() == hf.parse(data)["frames"]


@pytest.mark.parametrize("n,h,w,depth", [(3, 64, 64, 2), (36, 1024, 1024, 2), (11, 48, 80, 4)])
def test_store_frames_match_device(holo, tmp_path, n, h, w, depth):
    print(f"store n={n} {h}x{w} depth={depth}", flush=True)
    frames = _frames(n, h, w, depth, seed=3 * n)
    dev = torch.from_numpy(frames.view(np.uint8).reshape(n, -1).copy()).cuda()
    path = str(tmp_path / "out.holo")
    fd = holo.FrameDescriptor(w, h, depth)
    k = n // 3
    with holo.OutputHoloFile(path, fd, n, hf.PROCESSED) as out:
        out.meta_data = {"info": {"contiguous": n}}
        out.write_header()
        out.write_frames_from_device(dev[:k], k, torch.cuda.current_stream())
        out.write_frames_from_device(dev[k:], n - k, torch.cuda.current_stream())
        out.write_footer()
    assert open(path, "rb").read() == hf.holo_bytes(frames, depth, hf.PROCESSED,
                                                    footer=holo.dump_footer({"info": {"contiguous": n}}))


def test_replay_recording_through_pipe(holo, tmp_path):
    """c1 geometry (256x256 u16, B=S=T=16, angular spectrum, STFT): three
    batches recorded with the settings in the footer, replayed from HBM
    through a pipe configured from that footer, the u16 outputs recorded to
    a second .holo file; the same frames pushed directly give the same
    outputs, and the recorded file holds them bit for bit."""
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import Pipe, synth_frames
    s = settings_for("c1")
    B, H, W = s.batch_size, s.height, s.width
    frames = synth_frames(synth_for("c1"), 0, 3 * B)
    src = str(tmp_path / "rec.holo")
    fd = holo.FrameDescriptor(W, H, s.depth)
    with holo.OutputHoloFile(src, fd, 3 * B, hf.RAW) as out:
        out.export_compute_settings(s, input_fps=1000, contiguous=3 * B)
        out.write_header()
        out.write_frame(frames)
        out.write_footer()

    print("replay: direct pipe", flush=True)
    direct = Pipe(s)
    expect = []
    for b in range(3):
        if direct.push(torch.from_numpy(frames[b * B:(b + 1) * B].view(np.int16)).cuda()):
            expect.append(direct.output_u16())
    direct.close()
    assert len(expect) == 3

    print("replay: from the recording", flush=True)
    with holo.InputHoloFile(src) as f:
        cs = f.import_compute_settings()
        info = f.import_info()
        s2, ignored = holo.compute_settings_from_json(cs, f.get_frame_descriptor(), info, base=s)
        assert ignored == []
        for name, _ in type(s)._fields_:
            if name != "camera_fps":
                assert getattr(s2, name) == getattr(s, name), name
        # a recording made without a camera rate stores the input rate as camera rate
        assert s2.camera_fps == 1000.0
        dev = f.load_to_device()
    pipe = Pipe(s2)
    dst = str(tmp_path / "out.holo")
    got = []
    with holo.OutputHoloFile(dst, holo.FrameDescriptor(W, H, 2), 3, hf.PROCESSED) as out:
        out.export_compute_settings(s2, input_fps=1000, contiguous=3)
        out.write_header()
        for b in range(3):
            if pipe.push(dev[b * B:(b + 1) * B]):
                pipe.sync()
                u, _, _ = pipe._outputs()
                out.write_frames_from_device(u, 1, pipe.stream.value)
                got.append(pipe.output_u16())
        out.write_footer()
    pipe.close()
    assert len(got) == 3
    for a, b in zip(got, expect):
        assert np.array_equal(a, b)
    p = hf.parse(open(dst, "rb").read())
    assert p["img_nb"] == 3 and p["data_type"] == hf.PROCESSED
    assert p["frames"] == np.stack(expect).astype(np.uint16).tobytes()
