"""The record path (Backend/sources/core/pipe.cc:565-646,
Backend/sources/thread/frame_record_worker.cc:93-218): raw batches, the u16
output image or the three moment frames of each cycle, copied into a ring
of record slots in HBM on the pipe's stream (the reference's record queue)
and written from there into a .holo file by a writer thread
(hv_holo_store_frames: pinned double buffer, file writes overlapping the
DMA), so the file I/O overlaps the next cycles.  The file and the moments
are native; this class keeps the reference's record bookkeeping (frame
count, offset, skip, frame-count correction, footer)."""
from __future__ import annotations

import ctypes
import queue
import threading
import time

import torch

from . import _lib
from .holo_file import MOMENTS, PROCESSED, RAW, FrameDescriptor, OutputHoloFile

MODES = {"raw": RAW, "hologram": PROCESSED, "moments": MOMENTS}


class Recorder:
    """Records `nb_frames_to_record` frames of `mode` from `pipe` (one lane)
    into `path`.
    Call after_push(batch, produced) after every Pipe.push; close() writes
    the frame count and the footer.  frame_offset frames are skipped first
    (RecordFrameOffset), then frame_skip frames after each one recorded
    (FrameSkip)."""

    def __init__(self, path: str, pipe, mode: str, nb_frames_to_record: int, frame_offset: int = 0,
                 frame_skip: int = 0, slots: int = 4):
        if mode not in MODES:
            raise ValueError(f"record mode {mode!r}: one of {sorted(MODES)}")
        if getattr(pipe, "lanes", 1) > 1:
            # slots are filled on the pipe's first stream, in cycle order
            raise ValueError("the record path needs a pipe with one lane")
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
        # record slots in HBM: one cycle's frames each; the writer thread
        # drains them in order on its own stream
        per_cycle = {"raw": s.batch_size, "hologram": 1, "moments": 3}[mode]
        self._slots = [torch.empty(per_cycle * self.frame_size, dtype=torch.uint8, device="cuda")
                       for _ in range(max(1, slots))]
        self._free = queue.Queue()
        for i in range(len(self._slots)):
            self._free.put(i)
        self._todo = queue.Queue()
        self._error = None
        ps = pipe.stream.value if pipe.stream is not None else None
        self._pipe_stream = torch.cuda.ExternalStream(ps) if ps else torch.cuda.default_stream()
        self._writer = threading.Thread(target=self._write_loop, daemon=True)
        self._writer.start()

    def _write_loop(self) -> None:
        stream = torch.cuda.Stream()
        while True:
            item = self._todo.get()
            if item is None:
                return
            slot, runs, event = item
            try:
                event.synchronize()
                base = self._slots[slot].data_ptr()
                for k, n in runs:
                    self.file.write_frames_from_device(base + k * self.frame_size, n, stream)
            except Exception as e:  # reported by close()
                self._error = e
            finally:
                self._free.put(slot)

    def done(self) -> bool:
        return self.recorded >= self.to_record

    def _select(self, n: int):
        """(first frame, count) runs of the n frames of this cycle that are
        recorded, after offset and skip (frame_record_worker.cc:141-160)."""
        runs, k = [], 0
        while k < n and not self.done():
            if self.to_skip > 0:
                step = min(self.to_skip, n - k)
                self.to_skip -= step
                k += step
                continue
            run = 1 if self.frame_skip else min(n - k, self.to_record - self.recorded)
            runs.append((k, run))
            self.recorded += run
            k += run
            self.to_skip = self.frame_skip
        return runs

    def _enqueue(self, src: int, n: int) -> None:
        runs = self._select(n)
        if not runs:
            return
        slot = self._free.get()  # waits while every slot is still being written
        from .pipe import copy_d2d
        copy_d2d(self._slots[slot], src, n * self.frame_size, self.pipe.stream)
        event = torch.cuda.Event()
        event.record(self._pipe_stream)
        self._todo.put((slot, runs, event))

    def after_push(self, batch, produced: bool) -> None:
        if self.t0 is None:
            self.t0 = time.perf_counter()
        if self.done():
            return
        if self.mode == "raw":
            ptr = batch.data_ptr() if hasattr(batch, "data_ptr") else int(batch)
            self._enqueue(ptr, self.pipe.settings.batch_size)
            return
        if not produced:
            return
        if self.mode == "hologram":
            u, _, _ = self.pipe._outputs()
            self._enqueue(u, 1)
        else:
            f = ctypes.c_void_p()
            _lib.call("hv_pipe_record_frames", self.pipe.h, ctypes.byref(f))
            self._enqueue(f.value, 3)

    def close(self, input_fps: float | None = None) -> int:
        """correct_number_of_frames, export_compute_settings, write_footer
        (frame_record_worker.cc:177-199); returns the frames recorded."""
        if self.file is None:
            return self.recorded
        self._todo.put(None)
        self._writer.join()
        if self._error is not None:
            raise self._error
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
