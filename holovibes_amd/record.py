"""The record path (Backend/sources/core/pipe.cc:565-646,
Backend/sources/thread/frame_record_worker.cc:93-218): raw batches, the u16
output image or the three moment frames of each cycle, written straight
from HBM into a .holo file (hv_holo_store_frames: pinned double buffer,
file writes overlapping the DMA).  The file and the moments are native;
this class keeps the reference's record bookkeeping (frame count, offset,
skip, frame-count correction, footer)."""
from __future__ import annotations

import ctypes
import time

from . import _lib
from .holo_file import MOMENTS, PROCESSED, RAW, FrameDescriptor, OutputHoloFile

MODES = {"raw": RAW, "hologram": PROCESSED, "moments": MOMENTS}


class Recorder:
    """Records `nb_frames_to_record` frames of `mode` from `pipe` into `path`.
    Call after_push(batch, produced) after every Pipe.push; close() writes
    the frame count and the footer.  frame_offset frames are skipped first
    (RecordFrameOffset), then frame_skip frames after each one recorded
    (FrameSkip)."""

    def __init__(self, path: str, pipe, mode: str, nb_frames_to_record: int, frame_offset: int = 0,
                 frame_skip: int = 0):
        if mode not in MODES:
            raise ValueError(f"record mode {mode!r}: one of {sorted(MODES)}")
        s = pipe.settings
        self.pipe, self.mode = pipe, mode
        depth = {"raw": s.depth, "hologram": 2, "moments": 4}[mode]
        self.fd = FrameDescriptor(s.width, s.height, depth)
        self.frame_size = self.fd.get_frame_size()
        self.to_record = nb_frames_to_record
        self.to_skip = frame_offset
        self.frame_skip = frame_skip
        self.recorded = 0
        if mode == "moments":
            _lib.call("hv_pipe_set_record_moments", pipe.h, 1)
        self.file = OutputHoloFile(path, self.fd, nb_frames_to_record, MODES[mode])
        self.file.write_header()
        self.t0 = None

    def done(self) -> bool:
        return self.recorded >= self.to_record

    def _frames(self, base: int, n: int) -> None:
        """Frames base, base + frame_size, ... of a device buffer, after offset
        and skip; consecutive frames go to the file in one write."""
        k = 0
        while k < n and not self.done():
            if self.to_skip > 0:
                step = min(self.to_skip, n - k)
                self.to_skip -= step
                k += step
                continue
            run = 1 if self.frame_skip else min(n - k, self.to_record - self.recorded)
            self.file.write_frames_from_device(base + k * self.frame_size, run, self.pipe.stream.value)
            self.recorded += run
            k += run
            self.to_skip = self.frame_skip

    def after_push(self, batch, produced: bool) -> None:
        if self.t0 is None:
            self.t0 = time.perf_counter()
        if self.done():
            return
        if self.mode == "raw":
            ptr = batch.data_ptr() if hasattr(batch, "data_ptr") else int(batch)
            self.pipe.sync()
            self._frames(ptr, self.pipe.settings.batch_size)
            return
        if not produced:
            return
        self.pipe.sync()
        if self.mode == "hologram":
            u, _, _ = self.pipe._outputs()
            self._frames(u, 1)
        else:
            f = ctypes.c_void_p()
            _lib.call("hv_pipe_record_frames", self.pipe.h, ctypes.byref(f))
            self._frames(f.value, 3)

    def close(self, input_fps: float | None = None) -> int:
        """correct_number_of_frames, export_compute_settings, write_footer
        (frame_record_worker.cc:177-199); returns the frames recorded."""
        if self.file is None:
            return self.recorded
        self.file.correct_number_of_frames(self.recorded)
        if input_fps is None:
            el = time.perf_counter() - self.t0 if self.t0 else 0.0
            input_fps = self.recorded / el if el > 0 else 0.0
        self.file.export_compute_settings(self.pipe.settings, int(input_fps / (self.frame_skip + 1)), self.recorded)
        self.file.write_footer()
        self.file.close()
        self.file = None
        if self.mode == "moments":
            _lib.call("hv_pipe_set_record_moments", self.pipe.h, 0)
        return self.recorded
