"""Input side on the CPU: the unpack restatement against the definition of a
packed stream and hand-computed words, the queue model against the
reference's queue tests, and the C-ABI's argument checks (made before any
GPU call)."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import input_side as ins

HERE = os.path.dirname(os.path.abspath(__file__))


def kats():
    return json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))["batch_input_queue"]


@pytest.mark.parametrize("bits", [12, 10])
def test_unpack_oracle_matches_definition(bits):
    rng = np.random.default_rng(bits)
    v = rng.integers(0, 1 << bits, 4096).astype(np.uint32)
    v[:4] = [0, (1 << bits) - 1, 1, 1 << (bits - 1)]
    packed = ins.pack_msb_first(v, bits)
    assert packed.size == v.size * bits // 8
    out = (ins.unpack12 if bits == 12 else ins.unpack10)(packed, v.size)
    assert np.array_equal(out, (v << (16 - bits)).astype(np.uint16))


def test_unpack_oracle_hand_words():
    # 12-bit: 0xABC, 0xDEF packed as AB CD EF
    assert ins.unpack12(np.array([0xAB, 0xCD, 0xEF], np.uint8), 2).tolist() == [0xABC0, 0xDEF0]
    # 10-bit: 0x3FF, 0x000, 0x2AA, 0x155 -> 11111111 11000000 00001010 10101001 01010101
    words = ins.unpack10(np.array([0xFF, 0xC0, 0x0A, 0xA9, 0x55], np.uint8), 4).tolist()
    assert words == [0xFFC0, 0x0000, 0xAA80, 0x5540]


@pytest.mark.parametrize("case", kats(), ids=lambda c: c["name"])
def test_queue_model_reference_kats(case):
    fs = case["width"] * case["height"]
    q = ins.BatchInputQueueModel(case["capacity"], case["batch"], fs)
    frames = [ins.kat_frame(t, fs) for t in case["frames"]]
    for op in case["ops"]:
        if op[0] == "enq":
            q.enqueue(frames[op[1]])
        elif op[0] == "deq":
            got = q.dequeue()
            if op[1] is not None:
                assert got == [ins.kat_frame(t, fs) for t in op[1]]
        elif op[0] == "resize":
            q.resize(op[1])
        elif op[0] == "elt":
            assert q.data[op[1]] == ins.kat_frame(op[2], fs)
        if op[-1] is not None:
            assert q.size == op[-1], op


def test_unpack_argument_checks():
    from holovibes_amd import _lib
    lib = _lib.load()
    dummy = ctypes.c_void_p(0x1000)  # never dereferenced: the checks come first
    assert lib.hv_unpack_12_to_16bit(dummy, 3, dummy, 10, None) == _lib.HV_ERR_INVALID_ARG
    assert lib.hv_unpack_12_to_16bit(dummy, 8, dummy, 11, None) == _lib.HV_ERR_INVALID_ARG
    assert lib.hv_unpack_10_to_16bit(dummy, 6, dummy, 100, None) == _lib.HV_ERR_INVALID_ARG
    assert lib.hv_unpack_10_to_16bit(dummy, 8, dummy, 9, None) == _lib.HV_ERR_INVALID_ARG
    assert lib.hv_unpack_12_to_16bit(None, 8, dummy, 12, None) == _lib.HV_ERR_INVALID_ARG


def test_queue_argument_checks():
    from holovibes_amd import _lib
    lib = _lib.load()
    q = ctypes.c_void_p()
    assert lib.hv_input_queue_create(4, 0, 8, 8, 2, ctypes.byref(q)) == _lib.HV_ERR_INVALID_ARG
    assert lib.hv_input_queue_create(3, 4, 8, 8, 2, ctypes.byref(q)) == _lib.HV_ERR_INVALID_ARG
    assert b"at least a frame" in lib.hv_last_error()
    assert lib.hv_input_queue_create(4, 2, 0, 8, 2, ctypes.byref(q)) == _lib.HV_ERR_INVALID_ARG
