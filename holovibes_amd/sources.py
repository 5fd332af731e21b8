"""Frame sources (the reference's ICamera plugins, Camera/include/icamera.hh)
and the read worker that feeds the batch input queue
(Backend/sources/thread/camera_frame_read_worker.cc).  The built-in sources
and the worker thread are native (hv_source.cpp); `PyCamera` lets a camera
written in Python fill the same C function table."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (CAM_FD_FN, CAM_FPS_FN, CAM_FRAMES_FN, CAM_STATUS_FN, Camera, CapturedFrames, SynthParams)


class SyntheticCamera:
    """ring_frames generated frames in pinned memory, frames_per_call per get_frames."""

    def __init__(self, params: SynthParams, ring_frames: int, frames_per_call: int):
        self.table = Camera()
        _lib.call("hv_camera_synthetic", ctypes.byref(params), ring_frames, frames_per_call, ctypes.byref(self.table))

    def shutdown(self) -> None:
        if self.table.self:
            self.table.shutdown_camera(self.table.self)
            self.table.self = None


class HoloFileCamera:
    """A .holo recording replayed from first_frame to end_frame (0: the end), looping."""

    def __init__(self, path: str, frames_per_call: int, first_frame: int = 0, end_frame: int = 0):
        self.table = Camera()
        _lib.call("hv_camera_holo_file", path.encode(), first_frame, end_frame, frames_per_call,
                  ctypes.byref(self.table))

    def shutdown(self) -> None:
        if self.table.self:
            self.table.shutdown_camera(self.table.self)
            self.table.self = None


class PyCamera:
    """Base class of a camera written in Python: override get_frames (return
    a C-contiguous host array of whole frames, kept alive until the next
    call) and the frame descriptor."""

    def __init__(self, width: int, height: int, depth: int):
        self.fd = (width, height, depth)
        self._cur = None
        ok = lambda _self: 0  # noqa: E731
        self._fns = [CAM_STATUS_FN(lambda _s: self._wrap(self.init_camera)),
                     CAM_STATUS_FN(lambda _s: self._wrap(self.start_acquisition)),
                     CAM_STATUS_FN(lambda _s: self._wrap(self.stop_acquisition)),
                     CAM_STATUS_FN(ok), CAM_FRAMES_FN(self._get_frames), CAM_FD_FN(self._get_fd),
                     CAM_FPS_FN(lambda _s: -1)]
        self.table = Camera(None, *self._fns)

    @staticmethod
    def _wrap(fn) -> int:
        try:
            fn()
            return 0
        except Exception:
            return _lib.HV_ERR_INVALID_ARG

    def init_camera(self) -> None:
        pass

    def start_acquisition(self) -> None:
        pass

    def stop_acquisition(self) -> None:
        pass

    def get_frames(self) -> np.ndarray:
        raise NotImplementedError

    def _get_frames(self, _s, out) -> int:
        try:
            a = np.ascontiguousarray(self.get_frames())
            self._cur = a
            fs = self.fd[0] * self.fd[1] * self.fd[2]
            out.contents.region1 = a.ctypes.data
            out.contents.count1 = a.nbytes // fs
            out.contents.region2 = None
            out.contents.count2 = 0
            out.contents.on_gpu = 0
            return 0
        except Exception:
            return _lib.HV_ERR_INVALID_ARG

    def _get_fd(self, _s, w, h, d) -> int:
        w[0], h[0], d[0] = self.fd
        return 0

    def shutdown(self) -> None:
        pass


class FrameReadWorker:
    """A host thread moving a source's frames into a BatchInputQueue."""

    def __init__(self, camera, queue, fps_limit: float = 0.0):
        self.h = ctypes.c_void_p()
        self._camera = camera
        _lib.call("hv_reader_start", ctypes.byref(camera.table), queue.h, fps_limit, ctypes.byref(self.h))

    def frames(self) -> int:
        n = ctypes.c_uint64()
        _lib.call("hv_reader_frames", self.h, ctypes.byref(n))
        return n.value

    def stop(self) -> int:
        n = ctypes.c_uint64()
        if self.h:
            h, self.h = self.h, ctypes.c_void_p()
            _lib.call("hv_reader_stop", h, ctypes.byref(n))
        return n.value
