"""Input side on the GPU: packed 10/12-bit unpack against the oracle
(vector path, tails, unaligned buffers, a 16-frame 2048x2048 batch), the
HBM batch input queue against the reference's queue tests and the host
model (overwrite, resize, packed enqueue), and the queue feeding the pipe
in place, alone and from a producer thread."""
import ctypes
import json
import os
import threading

import numpy as np
import pytest
import torch

from oracle import input_side as ins

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _kats():
    return json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))["batch_input_queue"]


@pytest.mark.parametrize("bits", [12, 10])
@pytest.mark.parametrize("pixels,in_off,out_off", [(16, 0, 0), (8 * 1000 + 2, 0, 0), (16 * 999 + 12, 0, 0),
                                                   (4096, 1, 0), (4096, 0, 2), (2048 * 2048 * 16, 0, 0)])
def test_unpack_matches_oracle(bits, pixels, in_off, out_off):
    from holovibes_amd import steps
    unit = 2 if bits == 12 else 4
    pixels -= pixels % unit
    print(f"unpack bits={bits} pixels={pixels} in_off={in_off} out_off={out_off}", flush=True)
    rng = np.random.default_rng(pixels + bits)
    nbytes = pixels * bits // 8
    packed = rng.integers(0, 256, nbytes + in_off, dtype=np.uint8)
    dev_in = torch.from_numpy(packed).cuda()[in_off:]
    out_full = torch.full((pixels + out_off,), -1, dtype=torch.int16, device="cuda")
    out = out_full[out_off:]
    steps.unpack(dev_in, pixels, bits, out=out)
    got = out.cpu().numpy().view(np.uint16)
    expect = (ins.unpack12 if bits == 12 else ins.unpack10)(packed[in_off:], pixels)
    assert np.array_equal(got, expect)
    if out_off:
        assert int(out_full[0]) == -1


def _q(case):
    from holovibes_amd.input_queue import BatchInputQueue
    return BatchInputQueue(case["capacity"], case["batch"], case["width"], case["height"], 1)


@pytest.mark.parametrize("case", _kats(), ids=lambda c: c["name"])
def test_queue_reference_kats(case):
    print(f"queue kat {case['name']}", flush=True)
    fs = case["width"] * case["height"]
    q = _q(case)
    frames = [np.frombuffer(ins.kat_frame(t, fs), np.uint8).copy() for t in case["frames"]]
    for op in case["ops"]:
        if op[0] == "enq":
            q.enqueue(frames[op[1]], 1)
            torch.cuda.synchronize()
        elif op[0] == "deq":
            got = q.dequeue()
            if op[1] is not None:
                assert [bytes(r) for r in got] == [ins.kat_frame(t, fs) for t in op[1]]
        elif op[0] == "resize":
            q.resize(op[1])
        elif op[0] == "elt":
            torch.cuda.synchronize()
            t = torch.empty(fs, dtype=torch.uint8, device="cuda")
            from holovibes_amd.pipe import copy_d2d
            copy_d2d(t, q.data_ptr() + op[1] * fs, fs, None)
            torch.cuda.synchronize()
            assert bytes(t.cpu().numpy()) == ins.kat_frame(op[2], fs)
        if op[-1] is not None:
            assert q.get_size() == op[-1], op
    q.close()


def test_queue_matches_model_random_ops():
    """A long random mix of chunked enqueues, dequeues and drops: sizes,
    overwrite flag and every dequeued batch as the host model says."""
    from holovibes_amd.input_queue import BatchInputQueue
    rng = np.random.default_rng(5)
    W, H, B, cap = 64, 32, 4, 22  # 5 batches of 4 frames
    fs = W * H * 2
    q = BatchInputQueue(cap, B, W, H, 2)
    m = ins.BatchInputQueueModel(cap, B, fs)
    stamp = 0
    for step in range(300):
        r = rng.random()
        if r < 0.6:
            n = int(rng.integers(1, 11))
            fr = rng.integers(0, 65536, (n, H, W), dtype=np.uint16)
            fr[:, 0, 0] = np.arange(stamp, stamp + n) % 65536
            stamp += n
            q.enqueue(fr, n)
            for k in range(n):
                m.enqueue(fr[k].tobytes())
        elif r < 0.9:
            if m.size == 0 or m.producer_holds_start():
                assert q.front() is None
                continue
            got = q.dequeue()
            exp = m.dequeue()
            assert [bytes(x) for x in got] == exp, step
        else:
            ok = q.pop()
            assert ok == (m.size == 0 or not m.producer_holds_start())
            if m.size and ok:
                m.dequeue()
        st = q.state()
        assert (st["size"], st["start_index"], st["has_overwritten"]) == (m.size, m.start, m.has_overwritten), step
    q.close()


@pytest.mark.parametrize("bits,kind", [(12, "host"), (12, "device"), (10, "host")])
def test_packed_enqueue(bits, kind):
    """Packed frames enqueued in uneven chunks land unpacked in the queue's
    slots, batch by batch (host staging and device sources)."""
    from holovibes_amd.input_queue import BatchInputQueue
    W, H, B = 256, 128, 4
    P = W * H
    rng = np.random.default_rng(bits)
    vals = rng.integers(0, 1 << bits, (12, P)).astype(np.uint32)
    packed = np.stack([ins.pack_msb_first(v, bits) for v in vals])
    q = BatchInputQueue(8, B, W, H, 2)
    src = torch.from_numpy(packed).cuda() if kind == "device" else packed
    for a, b in [(0, 3), (3, 4), (4, 11), (11, 12)]:
        q.enqueue_packed(src[a:b] if kind == "device" else packed[a:b].copy(), b - a, bits)
    # 12 frames in a queue of two batches: the first batch was overwritten
    st = q.state()
    assert st["size"] == 2 and st["has_overwritten"]
    expect = (vals << (16 - bits)).astype(np.uint16)
    assert np.array_equal(q.dequeue().view(np.uint16).reshape(B, P), expect[4:8])
    assert np.array_equal(q.dequeue().view(np.uint16).reshape(B, P), expect[8:12])
    q.close()


def _pipe_frames(n):
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    s = settings_for("c1")
    return s, synth_frames(synth_for("c1"), 0, n)


def test_queue_feeds_pipe_in_place():
    """c1 frames enqueued 5 at a time (not batch-aligned); each whole batch
    goes from the queue to the pipe in place; outputs equal the pipe fed
    the same batches directly."""
    from holovibes_amd.input_queue import BatchInputQueue
    from holovibes_amd.pipe import Pipe
    s, frames = _pipe_frames(64)
    B = s.batch_size
    direct = Pipe(s)
    expect = []
    for b in range(4):
        if direct.push(torch.from_numpy(frames[b * B:(b + 1) * B].view(np.int16)).cuda()):
            expect.append(direct.output_u16())
    direct.close()
    q = BatchInputQueue(3 * B, B, s.width, s.height, 2)
    pipe = Pipe(s)
    got = []
    for k in range(0, 64, 5):
        q.enqueue(frames[k:k + 5].copy(), min(5, 64 - k))
        while True:
            b = q.front()
            if b is None:
                break
            made = pipe.push(b)
            q.pop()  # the pipe runs on the current stream
            if made:
                got.append(pipe.output_u16())
    pipe.close()
    q.close()
    assert len(got) == len(expect) == 4
    for a, e in zip(got, expect):
        assert np.array_equal(a, e)


def test_producer_thread_and_consumer():
    """A producer thread enqueues 40 stamped batches from pinned memory on
    its own stream while the consumer takes batches in place on another:
    every batch taken is whole and in order, and taken + overwritten = 40."""
    from holovibes_amd.input_queue import BatchInputQueue
    W, H, B, N = 512, 512, 8, 40
    q = BatchInputQueue(4 * B, B, W, H, 2)
    host = torch.empty((N * B, H, W), dtype=torch.int16, pin_memory=True)
    for i in range(N * B):
        host[i].fill_(i % 30000)
    prod_stream = torch.cuda.Stream()
    cons_stream = torch.cuda.Stream()
    errors = []

    def producer():
        try:
            for k in range(N):
                q.enqueue(host[k * B:(k + 1) * B].numpy(), B, prod_stream, pinned=True)
        except Exception as e:  # surfaced by the assertion below
            errors.append(e)

    th = threading.Thread(target=producer)
    th.start()
    taken = []
    out = torch.empty((B, H, W), dtype=torch.int16, device="cuda")
    from holovibes_amd.pipe import copy_d2d
    while th.is_alive() or q.get_size():
        b = q.front(cons_stream, timeout_ms=50)
        if b is None:
            continue
        copy_d2d(out, b, B * H * W * 2, ctypes.c_void_p(cons_stream.cuda_stream))
        q.pop(cons_stream)
        cons_stream.synchronize()
        v = out[:, 0, 0].cpu().numpy()
        assert np.array_equal(v, (v[0] + np.arange(B)) % 30000), v
        assert bool((out == out[:, :1, :1]).all())
        taken.append(int(v[0]) // B)
    th.join()
    assert not errors, errors
    assert taken == sorted(taken) and len(set(taken)) == len(taken)
    st = q.state()
    assert st["size"] == 0
    q.close()
