"""The output queue and the serving loop on the GPU.

Output queue (gpu_output_queue, Backend/sources/data_struct/queue.cc:149-189,
385-427): every produced u16 image lands in slot k mod max_size; with more
images than slots the oldest are overwritten; dequeue copies the oldest
(wrapping around the end) -- each image identical to the one a pipe without
a queue shows after that cycle, also with several lanes and graph replay.

Serving (continuous batching): requests of four sources at the serving
mix's real sizes are submitted without draining; every request's image,
taken from its source's output queue by the index the poll reports, is
identical to the image of the same batch when that source runs alone.  A
full waiting queue sheds its oldest requests.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _batches(name, n, **over):
    from holovibes_amd.configs import settings_for, synth_for
    from holovibes_amd.pipe import synth_frames
    s = settings_for(name, **over)
    syn = dict(width=s.width, height=s.height, period=s.time_transformation_size)
    if "depth" in over:  # a u16 source of a u8 workload: the 12-bit generator
        syn.update(depth=over["depth"], bits=12, mean_counts=1024.0)
    sp = synth_for(name, **syn)
    B = s.batch_size
    host = [synth_frames(sp, k * B, B) for k in range(n)]
    dev = [torch.from_numpy(h.view(np.int16) if h.dtype == np.uint16 else h).cuda() for h in host]
    torch.cuda.synchronize()
    return s, dev


def _alone(s, dev):
    """Images of a pipe without an output queue, one push at a time: {index of the produced image: u16}."""
    from holovibes_amd.pipe import Pipe
    s0 = type(s).from_buffer_copy(s)
    s0.output_buffer_size = 0
    p = Pipe(s0, graph=True)
    out = []
    for d in dev:
        if p.push(d):
            out.append(p.output_u16())
    p.close()
    return out


@pytest.mark.parametrize("lanes", [1, 3])
def test_output_queue_keeps_the_last_images(lanes):
    from holovibes_amd.pipe import Pipe
    s, dev = _batches("c1", 12, width=128, height=128, output_buffer_size=5)
    ref = _alone(s, dev)
    assert len(ref) == 12
    p = Pipe(s, graph=True, lanes=lanes)
    for d in dev:  # no synchronisation between pushes
        assert p.push(d)
    p.sync()
    q = p.output_queue()
    assert (q["max_size"], q["size"], q["start_index"], q["has_overwritten"], q["produced"]) == (5, 5, 12 % 5, True, 12)
    for k in range(12):
        img = p.output_frame(k)
        if k < 12 - 5:
            assert img is None, k  # overwritten
        else:
            assert np.array_equal(img, ref[k]), k
    assert np.array_equal(p.output_u16(), ref[-1])
    # dequeue 4 of the 5 (slots 2, 3, 4, 0: wraps around the end), then the last
    got = p.output_dequeue(4)
    for j in range(4):
        assert np.array_equal(got[j], ref[7 + j]), j
    q = p.output_queue()
    assert (q["size"], q["start_index"]) == (1, 1)
    assert np.array_equal(p.output_dequeue(1)[0], ref[11])
    p.close()


def test_server_c4_mix_full_size_without_draining():
    """The serving mix's four sources at their real sizes (256^2 T = 16,
    512^2 T = 32, 512^2 PCA T = 32, 1024^2 T = 64 with B = 32): all requests
    submitted interleaved with no drain; each request's image (by the output
    index its poll result names) equals the alone run's image of that batch."""
    from holovibes_amd.configs import SERVING_BUDGET_BYTES, SERVING_MIX
    from holovibes_amd.pipe import Server
    counts = {"c1": 8, "metric": 6, "c2": 4, "c3": 5}
    srcs = []
    for name, over, _rate in SERVING_MIX:
        print(f"serving mix source {name} alone", flush=True)
        s, dev = _batches(name, counts[name], **over)
        srcs.append((name, s, dev, _alone(s, dev)))
    srv = Server(SERVING_BUDGET_BYTES)
    srv.configure(2, 64)
    ids = [srv.add_source(s) for _, s, _, _ in srcs]
    assert all(i is not None for i in ids)
    print("serving mix: submitting every request", flush=True)
    req = {}  # request id -> (source index, batch index)
    for k in range(max(counts.values())):
        for j, (name, s, dev, _) in enumerate(srcs):
            if k < len(dev):
                req[srv.submit(ids[j], dev[k].data_ptr(), 0)] = (j, k)
    srv.drain()
    res = srv.poll_requests()
    assert len(res) == len(req) and srv.stats() == {"shed": 0, "pending": 0}
    seen = {j: [] for j in range(len(srcs))}
    for r in res:
        j, k = req[r["request_id"]]
        assert r["source"] == ids[j]
        if r["status"] == 0:
            seen[j].append(r["output_index"])
            name, s, _, ref = srcs[j]
            img = srv.output_frame(ids[j], r["output_index"], s.height, s.width)
            assert np.array_equal(img, ref[r["output_index"]]), (name, k)
        else:
            assert r["status"] == 1  # the c3 source's first batch fills half of its window
    for j, (name, _, _, ref) in enumerate(srcs):
        assert sorted(seen[j]) == list(range(len(ref))), name
    srv.close()


def test_server_sheds_oldest_waiting_requests():
    """One slot and two waiting places: eight requests submitted while the
    first runs (a 2048^2, T = 128 cycle takes milliseconds; the eight
    submissions well under one) leave the first running, shed the five
    oldest waiting ones and run the two newest."""
    from holovibes_amd.pipe import Server
    s, dev = _batches("c5_stft", 1, output_buffer_size=8)
    srv = Server(1 << 40)
    srv.configure(1, 2)
    sid = srv.add_source(s)
    srv.submit(sid, dev[0].data_ptr(), 0)
    srv.drain()  # graphs captured, warm
    srv.poll_requests()
    ids = [srv.submit(sid, dev[0].data_ptr(), 0) for _ in range(8)]
    st = srv.stats()
    assert st["pending"] == 2 and st["shed"] == 5, st
    srv.drain()
    res = {r["request_id"]: r["status"] for r in srv.poll_requests()}
    assert [res[i] for i in ids] == [0, 2, 2, 2, 2, 2, 0, 0], res
    srv.close()
