"""CPU restatement of the .holo file format (test infrastructure only).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the native
reader/writer (hv_holo_*); the product package never imports it.

Format (Backend/includes/io/holo_file.hh:58-83, #pragma pack(2), 64 bytes):
  magic "HOLO" | u16 version | u16 bits_per_pixel | u32 width | u32 height |
  u32 img_nb | u64 total_data_size | u8 endianness | u8 data_type | 34 pad
then img_nb frames of width*height*bits_per_pixel/8 bytes, then an optional
JSON footer running to the end of the file.

Pinned by: the reference's own Python harness layout of the same header
(tests/holo.py:23-33, '=4sHHIIIQB' + padding), the sizes of the reference's
test recordings (64 + n*W*H*depth, no footer), and round trips.
"""
from __future__ import annotations

import json
import struct

import numpy as np

HEADER = struct.Struct("<4sHHIIIQBB34s")
assert HEADER.size == 64
RAW, PROCESSED, MOMENTS = 0, 1, 2


class OracleFormatError(ValueError):
    pass


def pack_header(width: int, height: int, depth: int, img_nb: int, data_type: int = PROCESSED, version: int = 7,
                endianness: int = 0) -> bytes:
    """OutputHoloFile's header (output_holo_file.cc:19-33): bits = depth*8,
    total = W*H*depth*img_nb, little endian, padding zeroed."""
    return HEADER.pack(b"HOLO", version, depth * 8, width, height, img_nb & 0xFFFFFFFF,
                       width * height * depth * img_nb, endianness, data_type, b"\0" * 34)


def holo_bytes(frames: np.ndarray, depth: int, data_type: int = PROCESSED, footer: dict | bytes | None = None,
               version: int = 7) -> bytes:
    """A whole file: header, frames (n, H, W) as little-endian bytes, footer."""
    n, h, w = frames.shape[0], frames.shape[1], frames.shape[2]
    body = np.ascontiguousarray(frames).tobytes()
    assert len(body) == n * h * w * depth
    if isinstance(footer, dict):
        footer = json.dumps(footer, sort_keys=True, separators=(",", ":")).encode()
    return pack_header(w, h, depth, n, data_type, version) + body + (footer or b"")


def parse(data: bytes) -> dict:
    """InputHoloFile's checks and reads (input_holo_file.cc:19-90, 92-133):
    short header or wrong magic -> 'invalid bytecode size'; total size not
    frame_size * img_nb -> 'invalid total data size'; data type RAW before
    version 7; footer = bytes after the frames."""
    if len(data) < 64 or data[:4] != b"HOLO":
        raise OracleFormatError("Invalid holo file: invalid bytecode size")
    magic, version, bpp, w, h, n, total, endian, dtype, _ = HEADER.unpack(data[:64])
    depth = bpp // 8
    frame_size = w * h * depth
    if total != frame_size * n:
        raise OracleFormatError("Invalid holo file: invalid total data size")
    avail = max(0, len(data) - 64)
    frames_in_file = min(n, avail // frame_size) if frame_size else 0
    end = 64 + total
    return {"version": version, "bits_per_pixel": bpp, "width": w, "height": h, "img_nb": n, "depth": depth,
            "total_data_size": total, "endianness": 1 if endian else 0,
            "data_type": dtype if version == 7 else RAW, "frame_size": frame_size,
            "frames_in_file": frames_in_file, "truncated": end > len(data),
            "frames": data[64:64 + frames_in_file * frame_size], "footer": data[end:] if end <= len(data) else b""}
