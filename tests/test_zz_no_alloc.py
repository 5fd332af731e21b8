"""hv_percentile_range_noalloc: the percentile step with its histograms in
the library's static device array instead of a caller workspace.  Through
the C-ABI, two calls in a row (the second on a smaller image, so stale
histogram counts of the first would show) give exactly the values of
hv_percentile_range with a workspace and of the oracle's order statistics."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call_noalloc(lib, img, step, lo, hi, clamp):
    out = torch.full((2,), -7.0, dtype=torch.float32, device="cuda")
    H, W = img.shape
    st = lib.hv_percentile_range_noalloc(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(img.data_ptr()), W, H, step,
                                         ctypes.c_float(lo), ctypes.c_float(hi), clamp, None)
    assert st == 0, lib.hv_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _call_workspace(lib, img, step, lo, hi, clamp):
    nbytes = ctypes.c_size_t()
    assert lib.hv_percentile_scratch_size(ctypes.c_size_t(img.numel()), ctypes.byref(nbytes)) == 0
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    out = torch.empty(2, dtype=torch.float32, device="cuda")
    H, W = img.shape
    st = lib.hv_percentile_range(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(img.data_ptr()), W, H, step,
                                 ctypes.c_float(lo), ctypes.c_float(hi), clamp, ctypes.c_void_p(ws.data_ptr()), None)
    assert st == 0, lib.hv_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_percentile_matches_workspace_entry_point(oracle):
    from holovibes_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(31)
    cases = [((1024, 2048), 1, 0.5, 99.5, 0),   # the larger image first
             ((96, 64), 2, 0.02, 99.98, 1)]     # then a smaller one, subsampled, clamped
    for (H, W), step, lo, hi, clamp in cases:
        print(f"noalloc percentile {W}x{H} step {step}", flush=True)
        host = (rng.standard_normal((H, W)) * 500 + 3).astype(np.float32)
        img = torch.from_numpy(host).cuda()
        torch.cuda.synchronize()
        got = _call_noalloc(lib, img, step, lo, hi, clamp)
        want = _call_workspace(lib, img, step, lo, hi, clamp)
        assert np.array_equal(got, want), (got, want)
        src = oracle.subsample(host, step) if step > 1 else host
        ref = oracle.percentile(src, [lo, hi])
        if clamp:
            ref = np.maximum(ref, 1.0)
        assert np.array_equal(got, ref), (got, ref)


def test_noalloc_percentile_refuses_oversized_input():
    """More pixels than the 32-bit histogram counts hold: refused on the host,
    before any device work (the pointers are never dereferenced)."""
    from holovibes_amd import _lib
    lib = _lib.load()
    fake = ctypes.c_void_p(256)
    st = lib.hv_percentile_range_noalloc(fake, fake, 65536, 65537, 1, ctypes.c_float(1.0), ctypes.c_float(99.0), 0,
                                         None)
    assert st == _lib.HV_ERR_INVALID_ARG
    assert b"2^32" in lib.hv_last_error()
