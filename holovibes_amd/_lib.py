"""ctypes binding of libholo_hip.so (the C-ABI declared in include/holo_hip.h).

The library is built in-tree by holovibes_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# HOLOVIBES_AMD_LIB selects another build of the library (e.g. an experiment variant)
LIB_PATH = os.environ.get("HOLOVIBES_AMD_LIB", os.path.join(HERE, "libholo_hip.so"))

HV_OK = 0
HV_ERR_INVALID_ARG = 1
HV_ERR_UNSUPPORTED = 2
HV_ERR_OUT_OF_MEMORY = 3
HV_ERR_HIP = 4
HV_ERR_BUDGET = 5
HV_ERR_NOT_READY = 6
HV_ERR_IO = 7
HV_ERR_FORMAT = 8

_STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "UNSUPPORTED", 3: "OUT_OF_MEMORY", 4: "HIP", 5: "BUDGET",
                 6: "NOT_READY", 7: "IO", 8: "FORMAT"}


class HoloError(RuntimeError):
    def __init__(self, status: int, fn: str, msg: str):
        super().__init__(f"{fn} -> HV_ERR_{_STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class HoloHeader(ctypes.Structure):
    """hv_holo_header: the 64-byte .holo header, packed on 2 bytes."""
    _pack_ = 2
    _fields_ = [("magic_number", ctypes.c_char * 4), ("version", ctypes.c_uint16), ("bits_per_pixel", ctypes.c_uint16),
                ("img_width", ctypes.c_uint32), ("img_height", ctypes.c_uint32), ("img_nb", ctypes.c_uint32),
                ("total_data_size", ctypes.c_uint64), ("endianness", ctypes.c_uint8), ("data_type", ctypes.c_uint8),
                ("padding", ctypes.c_char * 34)]


class HoloInfo(ctypes.Structure):
    """hv_holo_info: what the reader derives from a .holo header."""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("img_nb", ctypes.c_uint32), ("version", ctypes.c_uint16), ("bits_per_pixel", ctypes.c_uint16),
                ("endianness", ctypes.c_uint8), ("data_type", ctypes.c_uint8), ("truncated", ctypes.c_uint8),
                ("reserved", ctypes.c_uint8), ("frame_size", ctypes.c_uint64), ("total_data_size", ctypes.c_uint64),
                ("frames_in_file", ctypes.c_uint64), ("footer_offset", ctypes.c_uint64),
                ("footer_size", ctypes.c_uint64)]


class CapturedFrames(ctypes.Structure):
    """hv_captured_frames: up to two regions of frames from a source."""
    _fields_ = [("region1", ctypes.c_void_p), ("count1", ctypes.c_uint32), ("region2", ctypes.c_void_p),
                ("count2", ctypes.c_uint32), ("on_gpu", ctypes.c_int)]


CAM_STATUS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)
CAM_FRAMES_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(CapturedFrames))
CAM_FD_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32),
                             ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32))
CAM_FPS_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)


class Camera(ctypes.Structure):
    """hv_camera: the camera interface as a C function table."""
    _fields_ = [("self", ctypes.c_void_p), ("init_camera", CAM_STATUS_FN), ("start_acquisition", CAM_STATUS_FN),
                ("stop_acquisition", CAM_STATUS_FN), ("shutdown_camera", CAM_STATUS_FN),
                ("get_frames", CAM_FRAMES_FN), ("get_fd", CAM_FD_FN), ("get_camera_fps", CAM_FPS_FN)]


class SynthParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("bits", ctypes.c_uint32), ("period", ctypes.c_uint32), ("sigma", ctypes.c_float),
                ("tilt_x", ctypes.c_float), ("tilt_y", ctypes.c_float), ("quad_phase", ctypes.c_float),
                ("mean_counts", ctypes.c_float), ("n_discs", ctypes.c_uint32), ("disc_cx", ctypes.c_float * 4),
                ("disc_cy", ctypes.c_float * 4), ("disc_r", ctypes.c_float * 4), ("fd_lo", ctypes.c_int32 * 4),
                ("fd_hi", ctypes.c_int32 * 4), ("seed", ctypes.c_uint64)]


class RequestResult(ctypes.Structure):
    """hv_request_result: one finished or shed serving request."""
    _fields_ = [("request_id", ctypes.c_int64), ("source", ctypes.c_int32), ("status", ctypes.c_int32),
                ("latency_ns", ctypes.c_int64), ("output_index", ctypes.c_int64)]


class ComputeSettings(ctypes.Structure):
    """hv_compute_settings: the reference settings the path reads."""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("batch_size", ctypes.c_uint32), ("time_stride", ctypes.c_uint32),
                ("time_transformation_size", ctypes.c_uint32), ("space_transformation", ctypes.c_int32),
                ("time_transformation", ctypes.c_int32), ("img_type", ctypes.c_int32), ("lambda_", ctypes.c_float),
                ("z_distance", ctypes.c_float), ("pixel_size", ctypes.c_float), ("p_start", ctypes.c_uint32),
                ("p_width", ctypes.c_uint32), ("unwrap_2d", ctypes.c_int32), ("fft_shift", ctypes.c_int32),
                ("contrast_enabled", ctypes.c_int32), ("contrast_invert", ctypes.c_int32),
                ("contrast_mode", ctypes.c_int32), ("contrast_min", ctypes.c_float), ("contrast_max", ctypes.c_float),
                ("contrast_lower", ctypes.c_float), ("contrast_upper", ctypes.c_float),
                ("contrast_subsample", ctypes.c_uint32), ("camera_fps", ctypes.c_float),
                ("lens_precision", ctypes.c_int32), ("pca_precision", ctypes.c_int32),
                ("renorm_enabled", ctypes.c_int32), ("renorm_constant", ctypes.c_uint32),
                ("img_accu_level", ctypes.c_uint32), ("log_enabled", ctypes.c_int32),
                ("cuts_enabled", ctypes.c_int32), ("x_start", ctypes.c_uint32), ("x_width", ctypes.c_uint32),
                ("y_start", ctypes.c_uint32), ("y_width", ctypes.c_uint32),
                ("registration_enabled", ctypes.c_int32), ("registration_zone", ctypes.c_float),
                ("contrast_reticle_enabled", ctypes.c_int32), ("contrast_reticle_scale", ctypes.c_float),
                ("output_buffer_size", ctypes.c_uint32)]


P = ctypes.c_void_p
I = ctypes.c_int
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
F = ctypes.c_float
I64 = ctypes.c_int64

# name -> argument types (every hv_* function returns hv_status unless listed in _RESTYPES)
SIGNATURES = {
    "hv_version": [],
    "hv_last_error": [],
    "hv_device_count": [P],
    "hv_convert_to_complex": [P, P, SZ, I, I, P],
    "hv_angular_spectrum_lens": [P, I, I, F, F, F, F, I, P],
    "hv_shift_corners_c64": [P, I, I, I, P],
    "hv_shift_corners_f32": [P, I, I, I, P],
    "hv_fresnel_transform_lens": [P, I, I, I, F, F, F, I, P],
    "hv_angular_spectrum": [P, I, P, I, I, P],
    "hv_fresnel_transform": [P, I, P, I, I, P],
    "hv_fft2": [P, I, I, I, I, P],
    "hv_propagate_to_ring": [P, I, I, I, I, I, P, P, I, I, P],
    "hv_prepare_lens": [P, I, I, I, F, F, F, I, P],
    "hv_stft": [P, P, I, SZ, P],
    "hv_complex_to_modulus": [P, P, I, I, SZ, P],
    "hv_complex_to_squared_modulus": [P, P, I, I, SZ, P],
    "hv_complex_to_argument": [P, P, I, I, SZ, P],
    "hv_stft_image": [P, P, I, SZ, I, I, I, P, P],
    "hv_stft_moments": [P, P, P, P, I, SZ, I, I, F, P],
    "hv_stft_cuts": [P, P, P, I, I, I, I, I, I, I, I, P],
    "hv_pca_workspace_size": [I, SZ, I, P],
    "hv_pca_cov": [P, P, I, SZ, I, P, P],
    "hv_eig_workspace_size": [I, P],
    "hv_eig_hermitian_f64": [P, P, P, I, P, P],
    "hv_pca_image": [P, P, P, I, SZ, I, I, I, P, P],
    "hv_pca_moments": [P, P, P, I, SZ, I, I, I, F, P, P],
    "hv_unwrap_scratch_size": [I, I, P],
    "hv_unwrap_2d": [P, P, I, I, P, P],
    "hv_rescale_unwrap2d": [P, P, SZ, P, P],
    "hv_subsample_frame_f32": [P, P, I, I, I, P],
    "hv_minmax_init": [P, P],
    "hv_minmax": [P, P, SZ, P],
    "hv_minmax_decode": [P, P, P],
    "hv_percentile_scratch_size": [SZ, P],
    "hv_percentile_range": [P, P, I, I, I, F, F, I, P, P],
    "hv_percentile_range_reticle": [P, P, I, I, F, F, F, I, P, P],
    "hv_percentile_range_noalloc": [P, P, I, I, I, F, F, I, P],
    "hv_apply_contrast": [P, SZ, F, F, P],
    "hv_float_to_ushort": [P, P, SZ, P],
    "hv_contrast_to_ushort": [P, P, SZ, P, I, P],
    "hv_default_settings": [P],
    "hv_validate_settings": [P],
    "hv_pipe_create": [P, P, P],
    "hv_pipe_memory_bytes": [P, P],
    "hv_pipe_push_batch": [P, P, P],
    "hv_pipe_push_batches": [P, P, I, P],
    "hv_pipe_output": [P, P, P, P],
    "hv_pipe_contrast_range": [P, P],
    "hv_pipe_set_graph": [P, I],
    "hv_pipe_set_fused": [P, I],
    "hv_pipe_path": [P, P],
    "hv_pipe_set_lanes": [P, I],
    "hv_pipe_time_stage": [P, I, P, I, I, P],
    "hv_pipe_input_buffer": [P, P],
    "hv_pipe_ring": [P, P, P, P],
    "hv_pipe_stream": [P, P],
    "hv_pipe_set_timing": [P, I],
    "hv_pipe_stage_times": [P, P, P],
    "hv_pipe_sync": [P],
    "hv_pipe_cuts": [P, P, P],
    "hv_pipe_registration_shift": [P, P, P],
    "hv_pipe_set_record_moments": [P, I],
    "hv_pipe_record_frames": [P, P],
    "hv_pipe_destroy": [P],
    "hv_synth_frames": [P, U64, U32, P],
    "hv_server_create": [SZ, P],
    "hv_server_add_source": [P, P, P],
    "hv_server_remove_source": [P, I],
    "hv_server_submit": [P, I, P, I64],
    "hv_server_poll": [P, P, P, I, P],
    "hv_server_drain": [P],
    "hv_server_output": [P, I, P, P],
    "hv_server_memory_bytes": [P, P],
    "hv_server_destroy": [P],
    "hv_server_configure": [P, I, I],
    "hv_server_submit_request": [P, I, P, I64, P],
    "hv_server_poll_requests": [P, P, I, P],
    "hv_server_output_frame": [P, I, I64, P, P],
    "hv_server_stats": [P, P, P],
    "hv_pipe_output_queue": [P, P, P, P, P, P],
    "hv_pipe_output_dequeue": [P, P, I, U32, P],
    "hv_pipe_output_frame": [P, U64, P, P],
    "hv_pipe_output_count": [P, P],
    "hv_pipe_measure_latency": [P, P, I, I, ctypes.c_double, P, P],
    "hv_unpack_12_to_16bit": [P, SZ, P, SZ, P],
    "hv_unpack_10_to_16bit": [P, SZ, P, SZ, P],
    "hv_input_queue_create": [U32, U32, U32, U32, U32, P],
    "hv_input_queue_enqueue": [P, P, I, U32, P],
    "hv_input_queue_enqueue_packed": [P, P, I, I, U32, P],
    "hv_input_queue_front": [P, P, I, P],
    "hv_input_queue_pop": [P, P],
    "hv_input_queue_dequeue": [P, P, I, P],
    "hv_input_queue_state": [P, P, P, P, P, P],
    "hv_input_queue_data": [P, P],
    "hv_input_queue_resize": [P, U32],
    "hv_input_queue_reset": [P],
    "hv_input_queue_destroy": [P],
    "hv_camera_synthetic": [P, U32, U32, P],
    "hv_camera_holo_file": [ctypes.c_char_p, U64, U64, U32, P],
    "hv_reader_start": [P, P, F, P],
    "hv_reader_frames": [P, P],
    "hv_reader_stop": [P, P],
    "hv_holo_make_header": [U32, U32, U32, U64, I, P],
    "hv_holo_open": [ctypes.c_char_p, P],
    "hv_holo_get_info": [P, P],
    "hv_holo_get_header": [P, P],
    "hv_holo_read_frames": [P, U64, U64, P, P],
    "hv_holo_load_frames": [P, U64, U64, P, P, P],
    "hv_holo_read_footer": [P, P, SZ, P],
    "hv_holo_close": [P],
    "hv_holo_create": [ctypes.c_char_p, U32, U32, U32, U64, I, P],
    "hv_holo_write_header": [P],
    "hv_holo_write_frames": [P, P, U64],
    "hv_holo_store_frames": [P, P, U64, P],
    "hv_holo_write_footer": [P, P, SZ],
    "hv_holo_correct_number_of_frames": [P, U64],
    "hv_holo_writer_close": [P],
}
_RESTYPES = {"hv_version": ctypes.c_char_p, "hv_last_error": ctypes.c_char_p, "hv_default_settings": None}

_lib = None


def load():
    """Loads libholo_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m holovibes_amd.build` "
                              "(the HIP path has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def call(name: str, *args) -> None:
    """Calls hv_<name>, raising HoloError on a non-OK status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != HV_OK:
        raise HoloError(st, name, lib.hv_last_error().decode(errors="replace"))


def status(name: str, *args) -> int:
    """Calls hv_<name> and returns its raw status."""
    return getattr(load(), name)(*args)
