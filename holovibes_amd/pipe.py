"""Python face of the per-batch Pipe (Backend/sources/core/pipe.cc) and of
the multi-source server.  The orchestration itself is C++ (hv_pipe.cpp);
this module only owns handles and copies results into torch tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import ComputeSettings

_hip = None
_D2D, _D2H, _H2D = 3, 2, 1


def hip():
    """The HIP runtime (already loaded by torch) for plain copies."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return _hip


def copy_d2d(dst: torch.Tensor, src_ptr: int, nbytes: int, stream) -> None:
    e = hip().hipMemcpyAsync(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src_ptr), nbytes, _D2D, stream)
    if e != 0:
        raise RuntimeError(f"hipMemcpyAsync failed ({e})")


class Pipe:
    """One source's compute pipe.  `settings` is a ComputeSettings; the pipe
    runs on torch's current stream unless `own_stream` (always with `graph`:
    the legacy default stream cannot be captured)."""

    # cycle paths (hv_pipe_set_fused levels)
    THREE_PASS, ROW_STFT, BINS = 0, 1, 2

    def __init__(self, settings: ComputeSettings, own_stream: bool = False, graph: bool = False,
                 fused: int = 2, lanes: int = 1):
        self.settings = settings
        self.h = ctypes.c_void_p()
        own_stream = own_stream or graph or lanes > 1
        stream = None if own_stream else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.call("hv_pipe_create", ctypes.byref(settings), stream, ctypes.byref(self.h))
        st = ctypes.c_void_p()
        _lib.call("hv_pipe_stream", self.h, ctypes.byref(st))
        self.stream = st
        if graph:
            _lib.call("hv_pipe_set_graph", self.h, 1)
        # fused: True (every eligible path, the default), False (three passes) or a level
        level = self.BINS if fused is True else int(fused)
        if level != self.BINS:
            _lib.call("hv_pipe_set_fused", self.h, level)
        if lanes > 1:
            _lib.call("hv_pipe_set_lanes", self.h, lanes)
        self.lanes = lanes
        self.P = settings.width * settings.height

    @staticmethod
    def memory_bytes(settings: ComputeSettings) -> int:
        n = ctypes.c_size_t()
        _lib.call("hv_pipe_memory_bytes", ctypes.byref(settings), ctypes.byref(n))
        return n.value

    def input_buffer(self) -> int:
        p = ctypes.c_void_p()
        _lib.call("hv_pipe_input_buffer", self.h, ctypes.byref(p))
        return p.value

    def push(self, frames) -> bool:
        """frames: device tensor (batch_size, H, W) or a raw device pointer."""
        ptr = frames.data_ptr() if isinstance(frames, torch.Tensor) else int(frames)
        produced = ctypes.c_int()
        _lib.call("hv_pipe_push_batch", self.h, ctypes.c_void_p(ptr), ctypes.byref(produced))
        return bool(produced.value)

    @staticmethod
    def batch_pointers(batches) -> ctypes.Array:
        """Device pointers of a batch sequence, for push_many."""
        return (ctypes.c_void_p * len(batches))(
            *[b.data_ptr() if isinstance(b, torch.Tensor) else int(b) for b in batches])

    def push_many(self, pointers: ctypes.Array) -> int:
        """hv_pipe_push_batches: every batch of `pointers` (batch_pointers) in
        one native loop; returns the number of output images written."""
        produced = ctypes.c_int()
        _lib.call("hv_pipe_push_batches", self.h, ctypes.cast(pointers, ctypes.c_void_p), len(pointers),
                  ctypes.byref(produced))
        return produced.value

    @property
    def path(self) -> int:
        """Cycle path in use: THREE_PASS, ROW_STFT or BINS."""
        v = ctypes.c_int()
        _lib.call("hv_pipe_path", self.h, ctypes.byref(v))
        return v.value

    def set_timing(self, enabled: bool) -> None:
        _lib.call("hv_pipe_set_timing", self.h, 1 if enabled else 0)

    def stage_times(self):
        """Mean ms of (row FFT, column pass, inverse row, image, contrast + u16) and the cycle count."""
        ms = (ctypes.c_float * 5)()
        n = ctypes.c_int()
        _lib.call("hv_pipe_stage_times", self.h, ms, ctypes.byref(n))
        return list(ms), n.value

    def time_stage(self, stage: int, frames, iters: int = 50) -> float:
        """Mean ms of one launch of a bins-first stage (0 rows, 1 columns,
        3 image, 4 contrast), `iters` launches back to back; stage 0 cycles
        over `frames` (device tensors or pointers)."""
        ptrs = [f.data_ptr() if isinstance(f, torch.Tensor) else int(f) for f in frames]
        arr = (ctypes.c_void_p * max(1, len(ptrs)))(*ptrs)
        ms = ctypes.c_float()
        _lib.call("hv_pipe_time_stage", self.h, stage, arr, len(ptrs), iters, ctypes.byref(ms))
        return ms.value

    def measure_latency(self, frames, n_batches: int, period_ns: float):
        """Batch -> image latencies (ns) of n_batches pushed at a fixed period
        (hv_pipe_measure_latency), cycling over `frames`."""
        ptrs = [f.data_ptr() if isinstance(f, torch.Tensor) else int(f) for f in frames]
        arr = (ctypes.c_void_p * len(ptrs))(*ptrs)
        lat = (ctypes.c_int64 * n_batches)()
        n = ctypes.c_int()
        _lib.call("hv_pipe_measure_latency", self.h, arr, len(ptrs), n_batches, float(period_ns), lat, ctypes.byref(n))
        return list(lat[:n.value])

    def sync(self) -> None:
        _lib.call("hv_pipe_sync", self.h)

    def _outputs(self):
        u, c, i = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        st = _lib.status("hv_pipe_output", self.h, ctypes.byref(u), ctypes.byref(c), ctypes.byref(i))
        if st not in (_lib.HV_OK, _lib.HV_ERR_NOT_READY):
            _lib.call("hv_pipe_output", self.h, ctypes.byref(u), ctypes.byref(c), ctypes.byref(i))
        return u.value, c.value, i.value

    def output_u16(self) -> np.ndarray:
        self.sync()  # the latest cycle may run on another lane's stream
        u, _, _ = self._outputs()
        t = torch.empty((self.settings.height, self.settings.width), dtype=torch.int16, device="cuda")
        copy_d2d(t, u, self.P * 2, self.stream)
        self.sync()
        return t.cpu().numpy().view(np.uint16)

    def image(self) -> torch.Tensor:
        """Float image before contrast (gpu_postprocess_frame)."""
        self.sync()
        _, _, i = self._outputs()
        t = torch.empty((self.settings.height, self.settings.width), dtype=torch.float32, device="cuda")
        copy_d2d(t, i, self.P * 4, self.stream)
        self.sync()
        return t

    def contrast_image(self) -> torch.Tensor:
        self.sync()
        _, c, _ = self._outputs()
        t = torch.empty((self.settings.height, self.settings.width), dtype=torch.float32, device="cuda")
        copy_d2d(t, c, self.P * 4, self.stream)
        self.sync()
        return t

    def output_queue(self) -> dict:
        """The output queue's state (Queue getters): base pointer, max_size,
        start_index, size, has_overwritten, and the images produced so far."""
        b, m, a, n, ow, c = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int(), \
            ctypes.c_uint64()
        _lib.call("hv_pipe_output_queue", self.h, ctypes.byref(b), ctypes.byref(m), ctypes.byref(a), ctypes.byref(n),
                  ctypes.byref(ow))
        _lib.call("hv_pipe_output_count", self.h, ctypes.byref(c))
        return {"base": b.value, "max_size": m.value, "start_index": a.value, "size": n.value,
                "has_overwritten": bool(ow.value), "produced": c.value}

    def output_frame(self, index: int):
        """The u16 image of the index-th produced image while the queue holds
        it (numpy), else None."""
        fr = ctypes.c_void_p()
        st = _lib.status("hv_pipe_output_frame", self.h, index, None, ctypes.byref(fr))
        if st == _lib.HV_ERR_NOT_READY:
            return None
        if st != _lib.HV_OK:
            _lib.call("hv_pipe_output_frame", self.h, index, None, ctypes.byref(fr))
        self.sync()
        t = torch.empty((self.settings.height, self.settings.width), dtype=torch.int16, device="cuda")
        copy_d2d(t, fr.value, self.P * 2, self.stream)
        self.sync()
        return t.cpu().numpy().view(np.uint16)

    def output_dequeue(self, n: int = 1) -> np.ndarray:
        """Queue::dequeue of the n oldest images into host memory."""
        out = np.empty((n, self.settings.height, self.settings.width), np.uint16)
        _lib.call("hv_pipe_output_dequeue", self.h, ctypes.c_void_p(out.ctypes.data), _D2H, n, None)
        return out

    def contrast_range(self) -> np.ndarray:
        self.sync()
        r = ctypes.c_void_p()
        _lib.call("hv_pipe_contrast_range", self.h, ctypes.byref(r))
        t = torch.empty(2, dtype=torch.float32, device="cuda")
        copy_d2d(t, r.value, 8, self.stream)
        self.sync()
        return t.cpu().numpy()

    def ring(self):
        """(ring copy (T, H, W) complex64, start_index, size)."""
        r, a, b = ctypes.c_void_p(), ctypes.c_uint32(), ctypes.c_uint32()
        _lib.call("hv_pipe_ring", self.h, ctypes.byref(r), ctypes.byref(a), ctypes.byref(b))
        T = self.settings.time_transformation_size
        t = torch.empty((T, self.settings.height, self.settings.width), dtype=torch.complex64, device="cuda")
        copy_d2d(t, r.value, T * self.P * 8, self.stream)
        self.sync()
        return t, a.value, b.value

    def close(self) -> None:
        if self.h:
            _lib.call("hv_pipe_destroy", self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Several sources served concurrently; admission by an HBM budget."""

    def __init__(self, budget_bytes: int):
        self.h = ctypes.c_void_p()
        _lib.call("hv_server_create", budget_bytes, ctypes.byref(self.h))

    def add_source(self, settings: ComputeSettings):
        """Returns the source id, or None when the budget refuses it."""
        sid = ctypes.c_int()
        st = _lib.status("hv_server_add_source", self.h, ctypes.byref(settings), ctypes.byref(sid))
        if st == _lib.HV_ERR_BUDGET:
            return None
        if st != _lib.HV_OK:
            _lib.call("hv_server_add_source", self.h, ctypes.byref(settings), ctypes.byref(sid))
        return sid.value

    def remove_source(self, sid: int) -> None:
        _lib.call("hv_server_remove_source", self.h, sid)

    def configure(self, slots: int, pending: int) -> None:
        """Requests on the GPU at once per source, and requests waiting per source."""
        _lib.call("hv_server_configure", self.h, slots, pending)

    def submit(self, sid: int, frames_ptr: int, t_arrival_ns: int) -> int:
        """Submits a request; returns its id."""
        rid = ctypes.c_int64()
        _lib.call("hv_server_submit_request", self.h, sid, ctypes.c_void_p(frames_ptr), t_arrival_ns,
                  ctypes.byref(rid))
        return rid.value

    def poll_requests(self, max_n: int = 4096):
        """Finished and shed requests: dicts with request_id, source, status
        (DONE 0, NO_OUTPUT 1, SHED 2), latency_ns, output_index."""
        from ._lib import RequestResult
        buf = (RequestResult * max_n)()
        n = ctypes.c_int()
        _lib.call("hv_server_poll_requests", self.h, buf, max_n, ctypes.byref(n))
        return [{"request_id": r.request_id, "source": r.source, "status": r.status, "latency_ns": r.latency_ns,
                 "output_index": r.output_index} for r in buf[:n.value]]

    def output_frame(self, sid: int, index: int, height: int, width: int):
        """A request's u16 image from its source's output queue (numpy), or None once overwritten."""
        fr = ctypes.c_void_p()
        st = _lib.status("hv_server_output_frame", self.h, sid, index, None, ctypes.byref(fr))
        if st == _lib.HV_ERR_NOT_READY:
            return None
        if st != _lib.HV_OK:
            _lib.call("hv_server_output_frame", self.h, sid, index, None, ctypes.byref(fr))
        self.drain()
        t = torch.empty((height, width), dtype=torch.int16, device="cuda")
        copy_d2d(t, fr.value, height * width * 2, None)
        hip().hipStreamSynchronize(None)
        return t.cpu().numpy().view(np.uint16)

    def stats(self) -> dict:
        shed, pend = ctypes.c_uint64(), ctypes.c_uint32()
        _lib.call("hv_server_stats", self.h, ctypes.byref(shed), ctypes.byref(pend))
        return {"shed": shed.value, "pending": pend.value}

    def poll(self, max_n: int = 4096):
        srcs = (ctypes.c_int * max_n)()
        lat = (ctypes.c_int64 * max_n)()
        n = ctypes.c_int()
        _lib.call("hv_server_poll", self.h, srcs, lat, max_n, ctypes.byref(n))
        return [(srcs[i], lat[i]) for i in range(n.value)]

    def drain(self) -> None:
        _lib.call("hv_server_drain", self.h)

    def output_u16(self, sid: int, height: int, width: int) -> np.ndarray:
        # the source's streams do not synchronise with the null stream the copy runs on
        self.drain()
        u, f = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("hv_server_output", self.h, sid, ctypes.byref(u), ctypes.byref(f))
        t = torch.empty((height, width), dtype=torch.int16, device="cuda")
        copy_d2d(t, u.value, height * width * 2, None)
        hip().hipStreamSynchronize(None)
        return t.cpu().numpy().view(np.uint16)

    def memory_bytes(self) -> int:
        n = ctypes.c_size_t()
        _lib.call("hv_server_memory_bytes", self.h, ctypes.byref(n))
        return n.value

    def close(self) -> None:
        if self.h:
            _lib.call("hv_server_destroy", self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synth_frames(params, first_frame: int, count: int) -> np.ndarray:
    """Host frames from the seeded generator (count, H, W) uint8/uint16."""
    dt = np.uint8 if params.depth == 1 else np.uint16
    out = np.empty((count, params.height, params.width), dt)
    _lib.call("hv_synth_frames", ctypes.byref(params), first_frame, count, out.ctypes.data_as(ctypes.c_void_p))
    return out
