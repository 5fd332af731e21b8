"""Python face of the batch input queue (Backend/sources/data_struct/
batch_input_queue.cc) in HBM: the queue, its overwrite policy and the
in-place hand-off of the oldest batch to the pipe run in libholo_hip.so
(hv_input_queue_*); this module owns the handle."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

H2D, D2H, D2D = 1, 2, 3


def _stream(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream)) if stream else None


def _src(frames):
    """(address, memcpy kind, temporary host copy or None) of host frames
    (numpy) or device frames (tensor)."""
    if isinstance(frames, torch.Tensor):
        if frames.is_cuda:
            if not frames.is_contiguous():
                raise ValueError("device frames must be contiguous")
            return ctypes.c_void_p(frames.data_ptr()), D2D, None
        frames = frames.numpy()
    a = np.ascontiguousarray(frames)
    return ctypes.c_void_p(a.ctypes.data), H2D, (a if a is not frames else None)


class BatchInputQueue:
    """Frames of width x height x depth bytes, batch_size per batch, up to
    frame_capacity frames (rounded down to whole batches)."""

    def __init__(self, frame_capacity: int, batch_size: int, width: int, height: int, depth: int):
        self.width, self.height, self.depth = width, height, depth
        self.frame_size = width * height * depth
        self.h = ctypes.c_void_p()
        _lib.call("hv_input_queue_create", frame_capacity, batch_size, width, height, depth, ctypes.byref(self.h))
        self.batch_size = batch_size
        self._held = None

    def enqueue(self, frames, nb_frames: int | None = None, stream=None, pinned: bool = False) -> None:
        """Appends frames: a contiguous device tensor, or a host numpy array.
        A host copy may still be reading the array when the call returns, so
        the stream is synchronised unless `pinned` says the array is pinned
        memory that the caller keeps alive until the stream's copy has run."""
        ptr, kind, tmp = _src(frames)
        n = nb_frames if nb_frames is not None else _nbytes(frames) // self.frame_size
        _lib.call("hv_input_queue_enqueue", self.h, ptr, kind, n, _stream(stream))
        if kind == H2D and (tmp is not None or not pinned):
            _sync(stream)

    def enqueue_packed(self, packed, nb_frames: int, bits: int, stream=None, pinned: bool = False) -> None:
        """Appends nb_frames frames packed on `bits` bits, unpacked into the
        queue (host arrays as in enqueue)."""
        ptr, kind, tmp = _src(packed)
        _lib.call("hv_input_queue_enqueue_packed", self.h, ptr, kind, bits, nb_frames, _stream(stream))
        if kind == H2D and (tmp is not None or not pinned):
            _sync(stream)

    def front(self, stream=None, timeout_ms: int = 0):
        """Device address of the oldest whole batch (held until pop), or None."""
        b = ctypes.c_void_p()
        st = _lib.status("hv_input_queue_front", self.h, _stream(stream), timeout_ms, ctypes.byref(b))
        if st == _lib.HV_ERR_NOT_READY:
            return None
        if st != _lib.HV_OK:
            _lib.call("hv_input_queue_front", self.h, _stream(stream), timeout_ms, ctypes.byref(b))
        return b.value

    def pop(self, stream=None) -> bool:
        """Releases the held batch, or drops the oldest whole one; False when
        the producer is still refilling the oldest slot."""
        st = _lib.status("hv_input_queue_pop", self.h, _stream(stream))
        if st == _lib.HV_ERR_NOT_READY:
            return False
        if st != _lib.HV_OK:
            _lib.call("hv_input_queue_pop", self.h, _stream(stream))
        return True

    def dequeue(self, stream=None) -> np.ndarray:
        """The oldest batch copied to host memory (batch_size, frame bytes)."""
        out = np.empty((self.batch_size, self.frame_size), np.uint8)
        _lib.call("hv_input_queue_dequeue", self.h, ctypes.c_void_p(out.ctypes.data), D2H, _stream(stream))
        return out

    def state(self) -> dict:
        v = [ctypes.c_uint32() for _ in range(4)]
        ow = ctypes.c_int()
        _lib.call("hv_input_queue_state", self.h, *[ctypes.byref(x) for x in v], ctypes.byref(ow))
        return {"size": v[0].value, "max_size": v[1].value, "start_index": v[2].value,
                "curr_nb_frames": v[3].value, "has_overwritten": bool(ow.value)}

    def get_size(self) -> int:
        return self.state()["size"]

    def data_ptr(self) -> int:
        p = ctypes.c_void_p()
        _lib.call("hv_input_queue_data", self.h, ctypes.byref(p))
        return p.value

    def resize(self, batch_size: int) -> None:
        _lib.call("hv_input_queue_resize", self.h, batch_size)
        self.batch_size = batch_size

    def reset(self) -> None:
        _lib.call("hv_input_queue_reset", self.h)

    def close(self) -> None:
        if self.h:
            _lib.call("hv_input_queue_destroy", self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sync(stream) -> None:
    if hasattr(stream, "synchronize"):
        stream.synchronize()
    else:
        torch.cuda.synchronize()


def _nbytes(frames) -> int:
    if isinstance(frames, torch.Tensor):
        return frames.numel() * frames.element_size()
    return np.asarray(frames).nbytes
