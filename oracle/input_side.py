"""CPU restatement of the input side: packed-pixel unpack and the batch
input queue (test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's CPU-baseline
leg as the checker of hv_unpack_* and hv_input_queue_*; the product package
never imports it.

unpack12 / unpack10 restate Backend/cuda_sources/unpack.cu:4-32 byte for
byte.  Pinned by pack_msb_first, which builds a packed stream from the
definition of the format (the pixel values' bits concatenated, most
significant first), independently of the byte formulas: unpacking it must
give each value in the top bits of its 16-bit word.

BatchInputQueueModel restates Backend/sources/data_struct/batch_input_queue.cc
(create_queue, enqueue, dequeue, resize); pinned by the reference's own
queue tests (Backend/tests/test_batch_input_queue.cc), transcribed in
tests/golden/reference_kats.json.
"""
from __future__ import annotations

import numpy as np


def unpack12(packed: np.ndarray, pixels: int) -> np.ndarray:
    """unpack.cu:4-16: pixels (even) u16 from pixels * 3 / 2 bytes."""
    b = np.asarray(packed, np.uint8)[: pixels * 3 // 2].reshape(-1, 3).astype(np.uint32)
    out = np.empty((pixels // 2, 2), np.uint32)
    out[:, 0] = (b[:, 0] << 8) | (b[:, 1] & 0xF0)
    out[:, 1] = ((b[:, 1] & 0x0F) << 12) | (b[:, 2] << 4)
    return (out.reshape(-1) & 0xFFFF).astype(np.uint16)


def unpack10(packed: np.ndarray, pixels: int) -> np.ndarray:
    """unpack.cu:18-32: pixels (multiple of 4) u16 from pixels * 5 / 4 bytes."""
    b = np.asarray(packed, np.uint8)[: pixels * 5 // 4].reshape(-1, 5).astype(np.uint32)
    out = np.empty((pixels // 4, 4), np.uint32)
    out[:, 0] = (b[:, 0] << 8) | (b[:, 1] & 0xC0)
    out[:, 1] = ((b[:, 1] & 0x3F) << 10) | ((b[:, 2] & 0xF0) << 2)
    out[:, 2] = ((b[:, 2] & 0x0F) << 12) | ((b[:, 3] & 0xFC) << 4)
    out[:, 3] = ((b[:, 3] & 0x03) << 14) | (b[:, 4] << 6)
    return (out.reshape(-1) & 0xFFFF).astype(np.uint16)


def pack_msb_first(values: np.ndarray, bits: int) -> np.ndarray:
    """The packed stream by definition: each value's `bits` bits, most
    significant first, concatenated, then cut into bytes."""
    v = np.asarray(values, np.uint32)
    shifts = np.arange(bits - 1, -1, -1, dtype=np.uint32)
    bitstream = ((v[:, None] >> shifts[None, :]) & 1).astype(np.uint8).reshape(-1)
    assert bitstream.size % 8 == 0
    return np.packbits(bitstream)


class BatchInputQueueModel:
    """The queue's state and contents on the host, one frame a bytes object."""

    def __init__(self, capacity: int, batch_size: int, frame_size: int):
        self.capacity, self.frame_size = capacity, frame_size
        self.create_queue(batch_size)

    def create_queue(self, batch_size: int) -> None:
        # batch_input_queue.cc:33-59
        assert batch_size > 0
        self.batch_size = batch_size
        total = self.capacity - self.capacity % batch_size
        assert total > 0
        self.max_size = total // batch_size
        self.data = [b"\0" * self.frame_size] * total
        self.start = self.end = self.size = self.counter = 0
        self.has_overwritten = False

    def enqueue(self, frame: bytes) -> None:
        # batch_input_queue.cc:124-187, one frame
        self.data[self.end * self.batch_size + self.counter] = frame
        self.counter += 1
        if self.counter == self.batch_size:
            self.counter = 0
            self.end = (self.end + 1) % self.max_size
            if self.size == self.max_size:
                self.has_overwritten = True
                self.start = (self.start + 1) % self.max_size
            else:
                self.size += 1

    def producer_holds_start(self) -> bool:
        """The producer is refilling the oldest slot of a full queue: the
        reference's consumer waits on that batch's mutex until it completes
        (batch_input_queue.cc:189-195)."""
        return self.counter > 0 and self.end == self.start

    def dequeue(self) -> list[bytes]:
        # batch_input_queue.cc:189-222
        assert self.size > 0 and not self.producer_holds_start()
        b = self.data[self.start * self.batch_size:(self.start + 1) * self.batch_size]
        self.start = (self.start + 1) % self.max_size
        self.size -= 1
        return b

    def resize(self, batch_size: int) -> None:
        # batch_input_queue.cc:241-266
        self.create_queue(batch_size)


def kat_frame(text: str, frame_size: int) -> bytes:
    """A frame of the reference's queue tests: a C string padded with NUL."""
    raw = text.encode() + b"\0"
    return raw + b"\0" * (frame_size - len(raw))
