"""Benchmark of the per-batch compute path on one MI355X.

A step is one Pipe cycle (hv_pipe_push_batch) over one batch of synthetic
frames already resident in HBM (the timed batches pushed through one native
loop, hv_pipe_push_batches, as the compute worker runs its cycles): conversion + angular spectrum into the
time-window ring, STFT + image type, contrast, u16.  Default workload:
512 x 512 u16, B = S = T = 32 (see holovibes_amd/configs.py "metric").

Prints one JSON line: frames/s (whole job, one GPU), ms per step, p50/p99
batch -> image latency, the roofline of the dominant kernel (HIP events
around each stage, on the pipe's stream) and the CPU oracle baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")


def measured_traffic(workload: str, path: str, stage: str):
    """HBM bytes per launch of a stage's kernel (FETCH_SIZE + WRITE_SIZE from
    rocprofv3, calibrated on a device copy) as the committed profile holds
    them for this workload and path, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)[f"{workload}/{path}"]["stages"][stage]
    except (OSError, KeyError, ValueError):
        return None
    return rec["fetch_bytes"] + rec["write_bytes"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="metric")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="override a compute setting of the workload (e.g. pca_precision=0)")
    ap.add_argument("--latency-batches", type=int, default=10000,
                    help="batches of the latency-under-load measurement")
    ap.add_argument("--load", type=float, action="append", default=None,
                    help="offered load(s) of the latency measurement, fractions of the measured maximum "
                         "(default 0.5, 0.65, 0.8; the top-level p50 / p99 are the last one's)")
    ap.add_argument("--no-fp32-ratio", action="store_true",
                    help="skip the plain-FP32-path throughput measurements")
    ap.add_argument("--graph", action="store_true",
                    help="replay each cycle from a hipGraph capture (eager launches are faster here)")
    ap.add_argument("--no-graph", action="store_true", help="(default) eager launches")
    ap.add_argument("--input-mib", type=int, default=512,
                    help="device memory spread over distinct input batches cycled through by the timed steps")
    ap.add_argument("--lanes", type=int, default=3,
                    help="cycles in flight on the GPU at once (bins-first path only; 1 elsewhere)")
    ap.add_argument("--path", type=int, default=2,
                    help="highest cycle path allowed: 0 three passes + STFT, 1 fused row-IFFT + STFT, 2 bins-first")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--python-push", action="store_true",
                    help="time one Python call per batch (hv_pipe_push_batch) instead of the native batch loop")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run warmup + this many steps (for rocprofv3), print the JSON without extras")
    ap.add_argument("--stage-iters", type=int, default=50)
    ap.add_argument("--serving", action="store_true",
                    help="serving mix instead: four sources, Poisson batch arrivals, HBM-budget admission")
    ap.add_argument("--seconds", type=float, default=20.0, help="serving run length")
    ap.add_argument("--holo-io", action="store_true",
                    help=".holo recording I/O: file (page cache) -> HBM and HBM -> file throughput")
    ap.add_argument("--holo-frames", type=int, default=256, help="--holo-io: 2048x2048 u16 frames in the file")
    ap.add_argument("--holo-dir", default="/tmp", help="--holo-io: directory of the recording")
    ap.add_argument("--record", choices=["moments", "hologram", "raw"], default=None,
                    help="frames/s of the workload at one lane with and without recording into a .holo file")
    ap.add_argument("--input-side", action="store_true",
                    help="packed 12/10-bit unpack (HBM roofline) and input-queue enqueue throughput")
    ap.add_argument("--slots", type=int, default=2, help="serving: requests of one source on the GPU at once")
    ap.add_argument("--pending", type=int, default=16, help="serving: waiting requests per source before shedding")
    ap.add_argument("--rate-scale", type=float, default=1.0, help="serving: multiply every source's batch rate")
    ap.add_argument("--profile-stages", action="store_true",
                    help="with --profile-steps: after the steps, time the stages and print (for rocprofv3)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also copy this many MiB device-to-device once (a plain copy kernel of known size, "
                         "the reference for rocprofv3 FETCH_SIZE / WRITE_SIZE scaling)")
    return ap.parse_args()


def stage_bytes(s, path):
    """Algorithmic HBM bytes per launch of each stage (names, bytes): every
    stage reads its input once and writes its output once.  `path` is the
    pipe's cycle path (0 three passes + STFT, 1 inverse row pass fused with
    the STFT, 2 bins-first)."""
    P = s.width * s.height
    B, T, d = s.batch_size, s.time_transformation_size, s.depth
    nb = min(B, T)
    space = s.space_transformation
    contrast = ("contrast_u16", P * 4 + P * 2 + P * 4)  # image in, u16 out, contrasted float out
    if path == 2:
        W = s.width
        moments = s.img_type in (5, 6, 7)
        pmin = s.p_start
        if moments:
            K = min(pmin + s.p_width, T - 1) - pmin + 1
        else:
            K = (min(pmin + s.p_width, T) if s.p_width else pmin) - pmin + 1
        # the fused kernels hold an image row's K plane rows in LDS (W <= 1024, T <= 64);
        # otherwise the time DFT, the row FFTs and the inverse row FFTs are passes of their own
        pca_w = T * max(T, 8) * 8 if s.time_transformation == 2 else 0  # projection weights in LDS
        fused_rows = W in (64, 128, 256, 512, 1024) and T <= 64 and not (W == 1024 and T > 32) and \
            (K + 1) * (W + W // 8) * 8 + pca_w <= 160 * 1024
        fused_image = space == 2 and 64 <= W <= 1024 and not moments
        rows = T * P * d + K * P * 8 + (P * 8 if space == 1 else 0)  # raw window in, K planes out (+ chirp)
        if not fused_rows:
            rows += K * P * 16  # the planes' forward row FFT
        cols = K * P * 16 + (P * 8 if space == 2 else 0)
        image = K * P * 8 + P * 4
        if space == 2 and not fused_image:
            image += K * P * 16  # the planes' inverse row FFT
        split = "" if fused_rows and (fused_image or space == 1) else "_split"
        if s.time_transformation == 2:
            # PCA: the raw window read by the mean pass and by the covariance pass (+ the means), the
            # eigensolver (latency-bound, one workgroup), then the projection in place of the time DFT
            return [("pca_projection_row_fft" + split, rows),
                    ("column_fft_lens_ifft" if space == 2 else "column_fft", cols),
                    ("pca_cov_eig", 2 * T * P * d + 8 * P),
                    ("row_ifft_image" + split if space == 2 else "bins_image", image), contrast]
        return [("time_dft_row_fft" + split, rows), ("column_fft_lens_ifft" if space == 2 else "column_fft", cols),
                ("empty", 0), ("row_ifft_image" + split if space == 2 else "bins_image", image), contrast]
    if space == 2:  # angular spectrum: row FFT (+convert), column FFT x lens x IFFT, inverse row FFT
        st0 = nb * P * (d + 8)
        st1 = nb * P * 16 + P * 8
        st2 = 0 if path == 1 else nb * P * 16
    elif space == 1:  # Fresnel: row FFT (+convert, x chirp), column FFT
        st0 = nb * P * (d + 8) + P * 8
        st1 = nb * P * 16
        st2 = 0
    else:
        st0, st1, st2 = nb * P * (d + 8), 0, 0
    if s.time_transformation == 2:  # covariance read + projection read + image write
        st3 = 2 * T * P * 8 + P * 4
    else:
        st3 = T * P * 8 + P * 4
    t3 = "row_ifft_stft_image" if path == 1 else ["image", "stft_image", "pca_image"][s.time_transformation]
    return [("row_fft_convert", st0), ("column_fft_lens_ifft" if space == 2 else "column_fft", st1),
            ("empty" if path == 1 else "row_ifft_to_ring", st2), (t3, st3), contrast]


def image_error(s, batch, gpu_image):
    """Relative L2 error of the GPU image against the FP64 oracle pipe on the
    same batch (the window is full after one batch when B = S = T)."""
    import numpy as np
    from oracle import oracle as o
    o.build()
    pipe = o.OraclePipe(s, 64)
    ref = None
    for _ in range(max(1, s.time_transformation_size // s.batch_size)):
        ref = pipe.push(batch)
    if ref is None:
        return None
    r = ref["image"]
    r = r.astype(np.float64)  # squared norms of PCA images overflow float32
    return float(np.linalg.norm((gpu_image.astype(np.float64) - r).ravel()) / np.linalg.norm(r.ravel()))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(name, settings, batches, seconds, cores=None):
    from oracle import oracle as o
    o.build()
    cores = cores or host_cores()
    o.set_threads(cores)
    pipe = o.OraclePipe(settings, 32)
    B, S, T = settings.batch_size, settings.time_stride, settings.time_transformation_size
    k = 0
    while pipe.push(batches[k % len(batches)]) is None:  # fill the window
        k += 1
    n, frames = 0, 0
    t0 = time.perf_counter()
    while True:
        pipe.push(batches[k % len(batches)])
        k += 1
        n += 1
        frames += B
        el = time.perf_counter() - t0
        if el >= seconds or n >= 4096:
            break
    return {"value": frames / el, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{n} batches of {B} frames ({name}) through the FP32 C++ oracle pipe, "
                      f"{el:.1f} s on {cores} host threads", "cpu_model": cpu_model()}


def throughput(s, dev_batches, level, lanes, steps, warmup):
    """Frames/s of a pipe at cycle path `level` with `lanes` lanes (eager
    launches) over `steps` timed pushes after the window fill and warm-up."""
    import torch

    from holovibes_amd.pipe import Pipe
    p = Pipe(s, fused=level, lanes=lanes)
    n = len(dev_batches)
    k = 0
    while not p.push(dev_batches[k % n]):
        k += 1
    for _ in range(warmup):
        p.push(dev_batches[k % n])
        k += 1
    torch.cuda.synchronize()
    ptrs = Pipe.batch_pointers([dev_batches[(k + i) % n] for i in range(steps)])
    t0 = time.perf_counter()
    p.push_many(ptrs)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    path = p.path
    p.close()
    return steps * s.batch_size / el, path


def fp32_ratios(args, metric_value, metric_settings, metric_batches):
    """This build's configuration against the plain FP32 path on the same GPU
    in the same process: the reference's step order (three passes + time
    transform, no fusion), one lane, FP32 covariance -- for the metric
    workload and for c2 (Fresnel + PCA)."""
    import numpy as np
    import torch

    from holovibes_amd import configs
    from holovibes_amd.pipe import synth_frames
    out = {}
    plain, _ = throughput(metric_settings, metric_batches, 0, 1, args.steps, args.warmup)
    out["metric"] = {"value": metric_value / plain, "frames_per_s": metric_value, "plain_fp32_frames_per_s": plain}
    s2 = configs.settings_for("c2")
    sp = configs.synth_for("c2")
    # 64 device batches (512 MiB, more than the last-level cache) from 16 generated ones
    host = [synth_frames(sp, k * s2.batch_size, s2.batch_size) for k in range(16)]
    dev = [torch.from_numpy(h).cuda() for h in host for _ in range(4)]
    fast, _ = throughput(s2, dev, 2, args.lanes, args.steps, args.warmup)
    plain2, _ = throughput(configs.settings_for("c2", pca_precision=0), dev, 0, 1, args.steps, args.warmup)
    out["c2"] = {"value": fast / plain2, "frames_per_s": fast, "plain_fp32_frames_per_s": plain2,
                 "config": f"split BF16 covariance, {args.lanes} lanes vs FP32 covariance, 1 lane"}
    out["plain_fp32_path"] = "reference step order (path 0), one lane, eager launches, FP32 covariance"
    return out


def serving(args):
    """Four sources with Poisson batch arrivals served side by side by one
    hv_server (per-source pipes on their own streams, hipGraph replay);
    batch -> image latency from arrival to completion."""
    import numpy as np
    import torch

    from holovibes_amd import configs
    from holovibes_amd.pipe import Pipe, Server, synth_frames

    torch.cuda.set_device(0)
    srv = Server(configs.SERVING_BUDGET_BYTES)
    srv.configure(args.slots, args.pending)
    sources = []
    for i, (name, over, rate) in enumerate(configs.SERVING_MIX):
        s = configs.settings_for(name, **over)
        sp = configs.synth_for(name, seed_offset=i, depth=s.depth,
                               bits=12 if s.depth == 2 else 8, mean_counts=1024.0 if s.depth == 2 else 64.0)
        B = s.batch_size
        host = [synth_frames(sp, k * B, B) for k in range(2)]
        dev = [torch.from_numpy(h.view(np.int16) if h.dtype == np.uint16 else h).cuda() for h in host]
        sid = srv.add_source(s)
        if sid is None:
            raise RuntimeError(f"source {name} refused by the HBM budget")
        sources.append({"name": name, "sid": sid, "rate": rate * args.rate_scale, "B": B, "dev": dev,
                        "P": s.width * s.height, "T": s.time_transformation_size, "k": 0,
                        "bytes": Pipe.memory_bytes(s)})
    torch.cuda.synchronize()
    by_sid = {src["sid"]: src for src in sources}
    # warm-up: fill every window and capture each source's graphs
    for src in sources:
        for _ in range(4 * max(1, src["T"] // src["B"])):
            srv.submit(src["sid"], src["dev"][src["k"] % 2].data_ptr(), time.perf_counter_ns())
            src["k"] += 1
    srv.drain()
    srv.poll_requests()
    # Poisson arrivals for every source over the run
    rng = np.random.default_rng(4000)
    events = []
    for src in sources:
        t = 0.0
        while True:
            t += rng.exponential(1.0 / src["rate"])
            if t >= args.seconds:
                break
            events.append((t, src["sid"]))
    events.sort()
    lat = {src["sid"]: [] for src in sources}
    frames = {src["sid"]: 0 for src in sources}  # frames of requests that completed
    shed = {src["sid"]: 0 for src in sources}

    def collect():
        for r in srv.poll_requests():
            sid_r = r["source"]
            if r["status"] == 2:
                shed[sid_r] += 1
                continue
            frames[sid_r] += by_sid[sid_r]["B"]
            if r["status"] == 0:
                lat[sid_r].append(r["latency_ns"])

    t0 = time.perf_counter_ns()
    for t_ev, sid in events:
        target = t0 + int(t_ev * 1e9)
        while time.perf_counter_ns() < target:
            collect()
        src = by_sid[sid]
        srv.submit(sid, src["dev"][src["k"] % 2].data_ptr(), target)
        src["k"] += 1
        collect()
    srv.drain()
    collect()
    el = (time.perf_counter_ns() - t0) / 1e9
    per = []
    all_lat = []
    for src in sources:
        l = sorted(lat[src["sid"]])
        all_lat += l
        per.append({"source": src["name"], "batches_per_s": src["rate"], "batch": src["B"],
                    "images": len(l), "shed": shed[src["sid"]],
                    "p50_latency_ms": (l[len(l) // 2] / 1e6) if l else None,
                    "p99_latency_ms": (l[min(len(l) - 1, int(0.99 * len(l)))] / 1e6) if l else None,
                    "batch_period_ms": 1e3 / src["rate"], "hbm_bytes": src["bytes"]})
    all_lat.sort()
    offered = sum(src["B"] * src["rate"] for src in sources)
    line = {"metric": "serving mix: aggregate frames/s + p50/p99 batch->image latency",
            "value": sum(frames.values()) / el, "unit": "frames/s", "n_gpus": 1, "seconds": el,
            "offered_frames_per_s": offered, "higher_is_better": True,
            "data": "synthetic (seeded interferogram generator)",
            "p50_latency_ms": all_lat[len(all_lat) // 2] / 1e6 if all_lat else None,
            "p99_latency_ms": all_lat[min(len(all_lat) - 1, int(0.99 * len(all_lat)))] / 1e6 if all_lat else None,
            "requests_shed": sum(shed.values()), "request_slots_per_source": args.slots,
            "waiting_places_per_source": args.pending,
            "budget_bytes": configs.SERVING_BUDGET_BYTES, "used_bytes": srv.memory_bytes(),
            "sources": per, "config": {"workload": "serving_mix", "rate_scale": args.rate_scale}}
    srv.close()
    print(json.dumps(line))


def holo_io(args):
    """A 2048x2048 u16 recording (c5 frame size) streamed from the page cache
    into HBM (hv_holo_load_frames) and recorded from HBM into a file
    (hv_holo_store_frames); the bound is the host<->device link, measured
    here as a pinned hipMemcpy of the same size."""
    import ctypes

    import numpy as np
    import torch

    from holovibes_amd import holo_file as hfm
    from holovibes_amd.pipe import hip

    torch.cuda.set_device(0)
    n, H, W = args.holo_frames, 2048, 2048
    fsz = H * W * 2
    total = n * fsz
    path = os.path.join(args.holo_dir, f"hv_bench_{os.getpid()}.holo")
    out_path = path + ".out"
    rng = np.random.default_rng(7)
    batch = rng.integers(0, 4096, (16, H, W), dtype=np.uint16)
    with hfm.OutputHoloFile(path, hfm.FrameDescriptor(W, H, 2), n, hfm.RAW) as out:
        out.write_header()
        for k in range(0, n, 16):
            out.write_frame(batch[: min(16, n - k)])
    dev = torch.empty((n, fsz), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    reps = max(1, args.steps // 40)

    def timed(fn):
        fn()  # warm: page cache, pinned staging
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    f = hfm.InputHoloFile(path)

    def load():
        f.set_pos_to_frame(0)
        assert f.read_frames_to_device(dev, n, stream) == n

    def store():
        # a fresh file each time (freeing the previous one's pages is not timed)
        if os.path.exists(out_path):
            os.remove(out_path)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with hfm.OutputHoloFile(out_path, hfm.FrameDescriptor(W, H, 2), n, hfm.PROCESSED) as o:
            o.write_header()
            o.write_frames_from_device(dev, n, stream)
        return time.perf_counter() - t0

    t_load = timed(load)
    store()
    t_store = sum(store() for _ in range(reps)) / reps
    assert open(out_path, "rb").read(64 + 4096)[64:] == open(path, "rb").read(64 + 4096)[64:]
    f.close()
    # the link: one pinned copy of the same size each way
    pinned = torch.empty(total, dtype=torch.uint8, pin_memory=True)
    h = hip()

    def copy(kind):
        src, dst = (pinned, dev) if kind == 1 else (dev, pinned)
        e = h.hipMemcpyAsync(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), total, kind,
                             ctypes.c_void_p(stream.cuda_stream))
        assert e == 0

    t_h2d = timed(lambda: copy(1))
    t_d2h = timed(lambda: copy(2))
    del pinned
    # the oracle's read of the same file: one thread, Python buffered I/O
    t0 = time.perf_counter()
    with open(path, "rb") as fh:
        fh.seek(64)
        data = fh.read(total)
    t_cpu = time.perf_counter() - t0
    assert len(data) == total
    del data
    os.remove(path)
    os.remove(out_path)
    gb = total / 1e9
    line = {"metric": ".holo ingest: recording (page cache) -> HBM", "value": gb / t_load, "unit": "GB/s",
            "n_gpus": 1, "steps": reps, "higher_is_better": True, "data": "synthetic 12-bit frames",
            "ms_per_step": t_load * 1e3,
            "record_GBps": gb / t_store,
            "roofline": {"bound": "host-device link", "achieved": gb / t_load, "peak": gb / t_h2d, "unit": "GB/s",
                         "frac": t_h2d / t_load, "peak_source": "pinned hipMemcpy H2D of the same bytes",
                         "record_peak": gb / t_d2h, "record_frac": t_d2h / t_store},
            "cpu_baseline": {"value": gb / t_cpu, "unit": "GB/s", "cores": 1, "kind": "port",
                             "sample": f"one buffered read of the {gb:.2f} GB of frames into host memory"},
            "config": {"workload": "holo_io", "frames": n, "width": W, "height": H, "depth_bytes": 2,
                       "bytes": total}}
    print(json.dumps(line))


def input_side(args):
    """Unpack of a 128-frame 2048x2048 batch packed on 12 and 10 bits
    (algorithmic HBM bytes: bits/8 read + 2 written per pixel) and the batch
    input queue filled from pinned host memory with 16-bit frames and with
    12-bit packed frames (unpacked on arrival)."""
    import numpy as np
    import torch

    from holovibes_amd import steps
    from holovibes_amd.input_queue import BatchInputQueue
    from oracle import input_side as ins

    torch.cuda.set_device(0)
    F, H, W = 128, 2048, 2048
    P = F * H * W
    reps = max(5, args.steps // 10)
    res = {}
    for bits in (12, 10):
        packed = torch.randint(0, 256, (P * bits // 8,), dtype=torch.uint8, device="cuda")
        out = torch.empty(P, dtype=torch.int16, device="cuda")
        for _ in range(3):
            steps.unpack(packed, P, bits, out=out)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            steps.unpack(packed, P, bits, out=out)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1) / 1e3 / reps
        algo = P * (bits / 8 + 2)
        res[bits] = {"us_per_launch": t * 1e6, "GBps": algo / t / 1e9, "frac": algo / t / 8e12,
                     "bytes": int(algo), "frames_per_s": F / t}
        del packed, out
    # oracle unpack of a bounded sample, one host thread
    sample = np.random.default_rng(1).integers(0, 256, 8 * H * W * 3 // 2, dtype=np.uint8)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < min(args.cpu_seconds, 5.0):
        ins.unpack12(sample, 8 * H * W)
        n += 8
    cpu_el = time.perf_counter() - t0
    cpu_fps = n / cpu_el
    # queue fill from pinned host memory: 16-bit frames, then 12-bit packed frames
    B = 32
    q = BatchInputQueue(4 * B, B, W, H, 2)
    host16 = torch.empty((B, H, W), dtype=torch.int16, pin_memory=True)
    host12 = torch.empty((B * H * W * 3 // 2,), dtype=torch.uint8, pin_memory=True)
    stream = torch.cuda.current_stream()

    def fill(fn):
        for _ in range(4):
            fn()
            q.pop()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
            q.pop()
        torch.cuda.synchronize()
        return B * reps / (time.perf_counter() - t0)

    fps16 = fill(lambda: q.enqueue(host16.numpy(), B, stream, pinned=True))
    fps12 = fill(lambda: q.enqueue_packed(host12.numpy(), B, 12, stream, pinned=True))
    q.close()
    line = {"metric": "packed 12-bit unpack: 2048x2048 frames/s (128-frame batch)", "value": res[12]["frames_per_s"],
            "unit": "frames/s", "n_gpus": 1, "steps": reps, "higher_is_better": True, "dtype": "u16",
            "data": "synthetic packed bytes",
            "roofline": {"bound": "hbm", "achieved": res[12]["GBps"], "peak": 8000.0, "unit": "GB/s",
                         "frac": res[12]["frac"], "algorithmic_bytes": res[12]["bytes"]},
            "unpack10": res[10], "unpack12": res[12],
            "queue_fill_frames_per_s": {"u16_from_pinned_host": fps16, "packed12_from_pinned_host": fps12},
            "cpu_baseline": {"value": cpu_fps, "unit": "frames/s", "cores": 1, "kind": "port",
                             "sample": f"numpy restatement of the 12-bit unpack, {n} frames in 8-frame chunks, "
                                       f"{cpu_el:.1f} s"},
            "config": {"workload": "input_side", "frames": F, "width": W, "height": H}}
    print(json.dumps(line))


def record(args):
    """The workload at one lane, eager, with and without the record path
    (every cycle's frames written from HBM into a .holo file under
    --holo-dir): frames/s of both and the record's share of a cycle."""
    import numpy as np
    import torch

    from holovibes_amd import configs
    from holovibes_amd.pipe import Pipe, synth_frames
    from holovibes_amd.record import Recorder

    torch.cuda.set_device(0)
    s = configs.settings_for(args.workload)
    sp = configs.synth_for(args.workload)
    B = s.batch_size
    dev = [torch.from_numpy(h.view(np.int16) if h.dtype == np.uint16 else h).cuda()
           for h in (synth_frames(sp, k * B, B) for k in range(2))]
    steps = max(8, args.steps // 10)
    path = os.path.join(args.holo_dir, f"hv_record_{os.getpid()}.holo")

    def run(rec_mode):
        p = Pipe(s)
        rec = Recorder(path, p, rec_mode, 1 << 30) if rec_mode else None
        for k in range(args.warmup // 4 + 2):
            made = p.push(dev[k % 2])
            if rec:
                rec.after_push(dev[k % 2], made)
        p.sync()
        n0 = rec.recorded if rec else 0
        t0 = time.perf_counter()
        for k in range(steps):
            made = p.push(dev[k % 2])
            if rec:
                rec.after_push(dev[k % 2], made)
        p.sync()
        el = time.perf_counter() - t0
        n = rec.recorded - n0 if rec else 0
        if rec:
            rec.close(input_fps=1)
            os.remove(path)
        p.close()
        return B * steps / el, n, el

    base, _, _ = run(None)
    with_rec, frames, el = run(args.record)
    frame_bytes = {"moments": 4, "hologram": 2, "raw": s.depth}[args.record] * s.width * s.height
    line = {"metric": f"{args.workload} frames/s with the {args.record} record", "value": with_rec, "unit": "frames/s",
            "n_gpus": 1, "steps": steps, "higher_is_better": True, "data": "synthetic (seeded interferogram generator)",
            "without_record": base, "record_share": 1.0 - with_rec / base, "frames_recorded": frames,
            "record_GBps": frames * frame_bytes / el / 1e9,
            "config": {"workload": args.workload, "record": args.record, "lanes": 1}}
    print(json.dumps(line))


def main():
    args = parse()
    if args.serving:
        serving(args)
        return
    if args.record:
        record(args)
        return
    if args.input_side:
        input_side(args)
        return
    if args.holo_io:
        holo_io(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the pipe shards by source, "
              "one pipe per GPU", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from holovibes_amd import configs
    from holovibes_amd.pipe import Pipe, synth_frames

    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    name = args.workload
    overrides = {}
    for kv in args.set:
        key, val = kv.split("=", 1)
        overrides[key] = float(val) if "." in val or "e" in val.lower() else int(val)
    s = configs.settings_for(name, **overrides)
    # the source follows overridden frame sizes, depths and windows
    sp = configs.synth_for(name, width=s.width, height=s.height, depth=s.depth, period=s.time_transformation_size)
    B, S, T, P = s.batch_size, s.time_stride, s.time_transformation_size, s.width * s.height
    host_batches = [synth_frames(sp, k * B, B) for k in range(4)]
    to_dev = (lambda a: torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).cuda())
    # distinct device input buffers covering more than the 256 MB last-level
    # cache, so that every step reads its batch from HBM
    batch_bytes = B * P * s.depth
    nbatches = min(64, max(4, -(-args.input_mib * (1 << 20) // batch_bytes)))  # <= the pipe's graph cache
    dev_batches = [to_dev(host_batches[k % 4]) for k in range(nbatches)]
    torch.cuda.synchronize()
    if args.copy_probe_mb:
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    pipe_level = args.path
    pipe = Pipe(s, graph=args.graph and not args.no_graph, fused=pipe_level)
    # cycles side by side (own buffers when every batch replaces the window, else a shared ring)
    lanes = args.lanes
    if lanes > 1:
        pipe.close()
        pipe = Pipe(s, graph=args.graph and not args.no_graph, fused=pipe_level, lanes=lanes)
    k = 0
    while not pipe.push(dev_batches[k % nbatches]):  # fill the time window
        k += 1
    for _ in range(nbatches):  # one graph per input buffer, captured outside the timed region
        pipe.push(dev_batches[k % nbatches])
        k += 1
    for _ in range(args.warmup):
        pipe.push(dev_batches[k % nbatches])
        k += 1
    torch.cuda.synchronize()
    steps = args.profile_steps or args.steps
    # the timed batches go through one native loop (hv_pipe_push_batches, the
    # compute worker's cycle loop), or one Python call each with --python-push
    seq = [dev_batches[(k + i) % nbatches] for i in range(steps)]
    ptrs = Pipe.batch_pointers(seq)
    t0 = time.perf_counter()
    if args.python_push:
        for b in seq:
            pipe.push(b)
    else:
        pipe.push_many(ptrs)
    k += steps
    t_host = time.perf_counter() - t0
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el / steps * 1e3
    frames_per_s = steps * B / el
    line = {"metric": "frames/s + p99 batch->image latency at 512^2 T=32; % HBM roofline; err vs FP64",
            "value": frames_per_s, "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": ms_per_step, "host_submit_ms_per_step": t_host / steps * 1e3,
            "host_loop": "python, one hv_pipe_push_batch call per batch" if args.python_push else
            "native: hv_pipe_push_batches over the timed batches", "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded interferogram generator)",
            "config": {"workload": name, "overrides": overrides, "width": s.width, "height": s.height, "depth_bytes": s.depth,
                       "batch": B, "stride": S, "T": T, "space": ["none", "fresnel", "angular_spectrum"][
                           s.space_transformation], "time": ["none", "stft", "pca"][s.time_transformation],
                       "graph": bool(args.graph and not args.no_graph), "lanes": lanes}}
    if args.profile_steps and not args.profile_stages:
        print(json.dumps(line))
        return

    path = pipe.path
    line["config"]["path"] = ["three_pass", "row_ifft_stft_fused", "bins_first"][path]
    sb = stage_bytes(s, path)
    if path == Pipe.BINS:
        # each stage launched `stage_iters` times back to back between two HIP
        # events on the pipe's stream (stage 0 cycling over the input batches)
        ms = [pipe.time_stage(i, dev_batches, args.stage_iters) if sb[i][1] > 0 else 0.0 for i in range(5)]
        timing = f"mean of {args.stage_iters} back-to-back launches per stage, HIP events on the pipe's stream"
    else:
        # HIP events between the stages of every cycle over a second timed region
        pipe.set_timing(True)
        for _ in range(steps):
            pipe.push(dev_batches[k % nbatches])
            k += 1
        ms, _ = pipe.stage_times()
        pipe.set_timing(False)
        timing = "HIP events between the stages of each cycle (includes the event overhead)"
    stages = []
    for (nm, b), t in zip(sb, ms):
        gbs = b / (t * 1e-3) / 1e9 if t > 0 and b > 0 else 0.0
        stages.append({"name": nm, "ms": t, "bytes": b, "GB/s": gbs, "frac": gbs / HBM_PEAK_GBS})
    # the dominant HBM-bound stage (the PCA eigensolver is latency-bound on one workgroup: reported, not ranked)
    dom = max((i for i in range(5) if sb[i][1] > 0 and sb[i][0] != "pca_cov_eig"), key=lambda i: ms[i])
    line["roofline"] = {"bound": "hbm", "kernel": sb[dom][0], "achieved": stages[dom]["GB/s"],
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": stages[dom]["frac"],
                        "traffic": measured_traffic(name, line["config"]["path"], sb[dom][0]),
                        "algorithmic_bytes": sb[dom][1], "timing": timing}
    if line["roofline"]["traffic"] is not None:
        line["roofline"]["traffic_source"] = "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch)"
    tot_b, tot_ms = sum(b for _, b in sb), sum(ms)
    line["pipeline_roofline"] = {"bytes_per_step": tot_b, "ms_per_step_stages": tot_ms,
                                 "GB/s": tot_b / (tot_ms * 1e-3) / 1e9, "frac": tot_b / (tot_ms * 1e-3) / 1e9 /
                                 HBM_PEAK_GBS,
                                 "step_GB/s": tot_b / (ms_per_step * 1e-3) / 1e9,
                                 "step_frac": tot_b / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS}
    line["stages"] = stages
    if args.profile_stages:
        print(json.dumps(line))
        return

    # batch -> image latency under load: batches arrive at a fixed period, a
    # fraction of the measured maximum rate; completions polled natively
    loads = args.load or [0.5, 0.65, 0.8]
    line["latency"] = {"batches_per_load": args.latency_batches, "loads": []}
    for load in loads:
        period_ms = ms_per_step / load
        lat = sorted(x / 1e6 for x in pipe.measure_latency(dev_batches, args.latency_batches, period_ms * 1e6))
        if not lat:
            continue
        p50, p99 = statistics.median(lat), lat[min(len(lat) - 1, int(0.99 * len(lat)))]
        line["latency"]["loads"].append({"offered_load": load, "frames_per_s": B / (period_ms * 1e-3),
                                         "batch_period_ms": period_ms, "p50_ms": p50, "p99_ms": p99,
                                         "max_ms": lat[-1], "images": len(lat), "p99_below_period": p99 < period_ms})
        line["p50_latency_ms"], line["p99_latency_ms"] = p50, p99
    ok = [x for x in line["latency"]["loads"] if x["p99_below_period"]]
    line["latency"]["highest_load_with_p99_below_period"] = max((x["offered_load"] for x in ok), default=None)
    line["batch_period_ms_at_max_rate"] = ms_per_step
    # unloaded: one batch at a time
    lat0 = []
    for _ in range(200):
        t1 = time.perf_counter_ns()
        made = pipe.push(dev_batches[k % nbatches])
        k += 1
        pipe.sync()
        if made:
            lat0.append((time.perf_counter_ns() - t1) / 1e6)
    lat0.sort()
    if lat0:
        line["latency"]["unloaded_p50_ms"] = statistics.median(lat0)
        line["latency"]["unloaded_p99_ms"] = lat0[min(len(lat0) - 1, int(0.99 * len(lat0)))]
    # image error against the FP64 oracle on one batch through a fresh pipe
    check = Pipe(s, fused=pipe_level)
    made = False
    for _ in range(max(1, T // B)):
        made = check.push(dev_batches[0])
    if made:
        err = image_error(s, host_batches[0], check.image().cpu().numpy())
        if err is not None:
            line["err_vs_fp64"] = {"image_rel_l2": err, "oracle": "C++ FP64 restatement of the reference pipe"}
    check.close()
    pipe.close()
    if not args.no_fp32_ratio and name == "metric" and not overrides:
        # vs_baseline stays null (no published figure); the same-GPU plain path is the denominator here
        line["vs_plain_fp32_path"] = fp32_ratios(args, frames_per_s, s, dev_batches)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(name, s, host_batches, args.cpu_seconds)
        line["cpu_baseline_1thread"] = cpu_baseline(name, s, host_batches, args.cpu_seconds / 2, cores=1)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
