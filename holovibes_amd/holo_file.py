""".holo recordings: the reference's InputHoloFile / OutputHoloFile
(Backend/sources/io/input_file/input_holo_file.cc:19-167,
Backend/sources/io/output_file/output_holo_file.cc:10-122) over the C-ABI
reader and writer of libholo_hip.so (hv_holo_*), plus the mapping between
the JSON footer's compute settings (Backend/includes/struct/*.hh, version
"v1") and hv_compute_settings.

The header checks, the frame I/O and the file <-> HBM streaming are native
(hv_holo.cpp); this module parses and writes the footer JSON, which the
reference does with nlohmann::json.
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ComputeSettings, HoloHeader, HoloInfo
from .footer_versions import FooterUpgradeError, upgrade_compute_settings, upgrade_footer
from .constants import (HV_CONTRAST_MANUAL, HV_CONTRAST_REFERENCE, HV_IMG_ARGUMENT, HV_IMG_MODULUS,
                        HV_IMG_MOMENTS_0, HV_IMG_MOMENTS_1, HV_IMG_MOMENTS_2, HV_IMG_SQUARED_MODULUS,
                        HV_SPACE_ANGULARSP, HV_SPACE_FRESNELTR, HV_SPACE_NONE, HV_TIME_NONE, HV_TIME_PCA,
                        HV_TIME_STFT)

# holovibes::RecordedDataType (Backend/includes/enum/enum_recorded_data_type.hh:15-21)
RAW, PROCESSED, MOMENTS = 0, 1, 2
HEADER_SIZE = 64
CURRENT_VERSION = 7  # Backend/includes/io/holo_file.hh:98

# JSON names of the enums (Backend/includes/enum/enum_space_transformation.hh:31-42,
# enum_time_transformation.hh:25-31, enum_img_type.hh:28-37, enum_computation.hh:23-26);
# the first name of each value is the one written, the others are read for compatibility
_SPACE_NAMES = {"NONE": HV_SPACE_NONE, "None": HV_SPACE_NONE, "FRESNELTR": HV_SPACE_FRESNELTR,
                "Fresnel Tr.": HV_SPACE_FRESNELTR, "FFT1": HV_SPACE_FRESNELTR, "1FFT": HV_SPACE_FRESNELTR,
                "ANGULARSP": HV_SPACE_ANGULARSP, "Angular SP.": HV_SPACE_ANGULARSP, "FFT2": HV_SPACE_ANGULARSP,
                "2FFT": HV_SPACE_ANGULARSP}
_TIME_NAMES = {"NONE": HV_TIME_NONE, "None": HV_TIME_NONE, "STFT": HV_TIME_STFT, "PCA": HV_TIME_PCA}
_IMG_NAMES = {"MODULUS": HV_IMG_MODULUS, "SQUARED_MODULUS": HV_IMG_SQUARED_MODULUS, "ARGUMENT": HV_IMG_ARGUMENT,
              "MOMENTS_0": HV_IMG_MOMENTS_0, "MOMENTS_1": HV_IMG_MOMENTS_1, "MOMENTS_2": HV_IMG_MOMENTS_2}
_SPACE_WRITE = {HV_SPACE_NONE: "NONE", HV_SPACE_FRESNELTR: "FRESNELTR", HV_SPACE_ANGULARSP: "ANGULARSP"}
_TIME_WRITE = {HV_TIME_NONE: "NONE", HV_TIME_STFT: "STFT", HV_TIME_PCA: "PCA"}
_IMG_WRITE = {v: k for k, v in _IMG_NAMES.items()}


class HoloFormatError(ValueError):
    """A .holo file or footer the reader refuses (the reference's FileException)."""


class UnsupportedSettings(ValueError):
    """Footer compute settings that select a computation outside this path."""


@dataclass
class FrameDescriptor:
    """camera::FrameDescriptor (Camera/include/frame_desc.hh:25-50)."""
    width: int
    height: int
    depth: int  # bytes per pixel
    byte_endian: int = 0  # 0 little, 1 big

    def get_frame_size(self) -> int:
        return self.width * self.height * self.depth

    def get_frame_res(self) -> int:
        return self.width * self.height


_NP_DTYPE = {1: np.uint8, 2: np.uint16, 4: np.float32, 8: np.float64}


def _frames_array(nbytes_per_frame: int, n: int, depth: int, height: int, width: int) -> np.ndarray:
    dt = _NP_DTYPE.get(depth)
    if dt is None:
        return np.empty((n, nbytes_per_frame), np.uint8)
    return np.empty((n, height, width), dt)


class InputHoloFile:
    """Reader of a .holo file (input_holo_file.cc:19-167): the header is
    checked at open, frames are read to host memory or streamed to HBM, the
    footer JSON gives the compute settings and the acquisition info."""

    def __init__(self, file_path: str):
        self.file_path = file_path
        self.h = ctypes.c_void_p()
        try:
            _lib.call("hv_holo_open", file_path.encode(), ctypes.byref(self.h))
        except _lib.HoloError as e:
            if e.status == _lib.HV_ERR_FORMAT:
                raise HoloFormatError(str(e)) from e
            raise
        info = HoloInfo()
        _lib.call("hv_holo_get_info", self.h, ctypes.byref(info))
        self.info = info
        self.fd = FrameDescriptor(info.width, info.height, info.depth, info.endianness)
        self.version = info.version
        self.data_type = info.data_type
        self.has_footer = info.footer_size > 0 or bool(info.truncated)
        self._pos = 0
        self.meta_data: dict = {}

    # -- frames --------------------------------------------------------------------------------
    def get_total_nb_frames(self) -> int:
        return self.info.img_nb

    def get_frame_descriptor(self) -> FrameDescriptor:
        return self.fd

    def header(self) -> HoloHeader:
        h = HoloHeader()
        _lib.call("hv_holo_get_header", self.h, ctypes.byref(h))
        return h

    def set_pos_to_frame(self, frame_id: int) -> None:
        """input_holo_file.cc:45-51."""
        if frame_id < 0 or frame_id > self.info.frames_in_file:
            raise HoloFormatError("Unable to seek the frame requested")
        self._pos = frame_id

    def read_frames(self, frames_to_read: int) -> np.ndarray:
        """Reads up to frames_to_read frames from the current position into
        host memory (input_frame_file.cc:7-17); the position advances."""
        out = _frames_array(self.info.frame_size, frames_to_read, self.fd.depth, self.fd.height, self.fd.width)
        n = ctypes.c_uint64()
        _lib.call("hv_holo_read_frames", self.h, self._pos, frames_to_read, out.ctypes.data, ctypes.byref(n))
        self._pos += n.value
        return out[: n.value]

    def read_frames_to_device(self, dst, frames_to_read: int, stream=None) -> int:
        """Streams up to frames_to_read frames from the current position into
        device memory (a torch tensor of at least that many frames, or a
        device address) -- file_frame_read_worker.cc:197-271.  Returns the
        number of frames loaded; the position advances."""
        ptr, nbytes = _device_ptr(dst)
        if nbytes is not None and nbytes < frames_to_read * self.info.frame_size:
            raise ValueError(f"destination holds {nbytes} bytes, {frames_to_read} frames need "
                             f"{frames_to_read * self.info.frame_size}")
        n = ctypes.c_uint64()
        _lib.call("hv_holo_load_frames", self.h, self._pos, frames_to_read, ctypes.c_void_p(ptr),
                  _stream_arg(stream), ctypes.byref(n))
        self._pos += n.value
        return n.value

    def load_to_device(self, first: int = 0, count: int | None = None, device: str = "cuda"):
        """The whole recording (or [first, first + count)) as one uint8 torch
        tensor in HBM, (frames, frame_size) -- FileLoadKind::GPU of the reference."""
        import torch
        count = self.info.frames_in_file - first if count is None else count
        t = torch.empty((count, self.info.frame_size), dtype=torch.uint8, device=device)
        self.set_pos_to_frame(first)
        n = self.read_frames_to_device(t, count, torch.cuda.current_stream())
        return t[:n]

    # -- footer --------------------------------------------------------------------------------
    def footer_bytes(self) -> bytes:
        size = ctypes.c_size_t()
        try:
            _lib.call("hv_holo_read_footer", self.h, None, 0, ctypes.byref(size))
            buf = ctypes.create_string_buffer(max(1, size.value))
            _lib.call("hv_holo_read_footer", self.h, buf, size.value, ctypes.byref(size))
        except _lib.HoloError as e:
            if e.status == _lib.HV_ERR_FORMAT:
                raise HoloFormatError(str(e)) from e
            raise
        return buf.raw[: size.value]

    def import_compute_settings(self) -> dict:
        """The footer's "compute_settings" object, {} without a footer
        (input_holo_file.cc:92-140).  A footer that does not parse is skipped,
        as in the reference.  Footers of header versions 2 to 4 are brought
        to the current layout first (footer_versions.upgrade_footer); a
        footer the upgrade cannot convert, or of a version above 7, gives {}
        as in the reference."""
        if not self.has_footer:
            return {}
        raw = self.footer_bytes()
        self.meta_data = {}
        if raw:
            try:
                self.meta_data = json.loads(raw.decode("utf-8"))
            except (UnicodeDecodeError, json.JSONDecodeError):
                self.meta_data = {}
        try:  # the footer brought to the version-7 layout (holo_file_converter.cc:54-124)
            self.meta_data = upgrade_footer(self.meta_data, self.version)
        except FooterUpgradeError:
            return {}  # the reference then imports no settings (input_holo_file.cc:135-136)
        cs = self.meta_data.get("compute_settings") if isinstance(self.meta_data, dict) else None
        return cs if isinstance(cs, dict) else {}

    def import_info(self) -> dict:
        """Acquisition info of the footer (input_holo_file.cc:142-165):
        pixel_size (um, square pixels), input_fps, camera_fps (input_fps when
        absent) and eye_type when present; {} without a usable "info"."""
        if not self.has_footer:
            return {}
        if not self.meta_data:
            self.import_compute_settings()
        try:
            info = self.meta_data["info"]
            out = {"pixel_size": float(info["pixel_pitch"]["x"]), "input_fps": info["input_fps"]}
            out["camera_fps"] = info["camera_fps"] if "camera_fps" in info else info["input_fps"]
            if "eye_type" in info:
                out["eye_type"] = info["eye_type"]
            return out
        except (KeyError, TypeError, ValueError):
            return {}

    def close(self) -> None:
        if self.h:
            _lib.call("hv_holo_close", self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OutputHoloFile:
    """Writer of a .holo file (output_holo_file.cc:10-122): header at open
    (written by write_header), frames from host or device memory, the JSON
    footer, and the frame-count correction of an interrupted record."""

    def __init__(self, file_path: str, fd: FrameDescriptor, img_nb: int, data_type: int = PROCESSED):
        self.file_path = file_path
        self.fd = fd
        self.h = ctypes.c_void_p()
        _lib.call("hv_holo_create", file_path.encode(), fd.width, fd.height, fd.depth, img_nb, data_type,
                  ctypes.byref(self.h))
        self.meta_data: dict = {}

    def export_compute_settings(self, settings: ComputeSettings, input_fps: int, contiguous: int,
                                camera_fps: float | None = None, eye_type: str = "NONE") -> None:
        """output_holo_file.cc:39-61: footer = {"compute_settings": ...,
        "info": {pixel_pitch, input_fps, camera_fps, eye_type, contiguous,
        holovibes_version}}.  camera_fps defaults to settings.camera_fps, or
        input_fps when that is 0."""
        cam = settings.camera_fps if camera_fps is None else camera_fps
        known = bool(cam)  # the reference's can_get_camera_fps()
        cam = _num(cam if known else input_fps)
        px = float(settings.pixel_size)
        info = {"pixel_pitch": {"x": px, "y": px}, "input_fps": cam if known else _num(input_fps),
                "camera_fps": cam, "eye_type": eye_type,
                "contiguous": int(contiguous), "holovibes_version": _lib.load().hv_version().decode()}
        self.meta_data = {"compute_settings": compute_settings_to_json(settings), "info": info}

    def write_header(self) -> None:
        _lib.call("hv_holo_write_header", self.h)

    def write_frame(self, frame) -> int:
        """Appends frames held in host memory (one frame or a stack)."""
        a = np.ascontiguousarray(frame)
        fs = self.fd.get_frame_size()
        if a.nbytes % fs:
            raise ValueError(f"{a.nbytes} bytes is not a whole number of {fs}-byte frames")
        _lib.call("hv_holo_write_frames", self.h, a.ctypes.data, a.nbytes // fs)
        return a.nbytes

    def write_frames_from_device(self, src, count: int, stream=None) -> None:
        """Appends count frames held in device memory (tensor or address)."""
        ptr, nbytes = _device_ptr(src)
        if nbytes is not None and nbytes < count * self.fd.get_frame_size():
            raise ValueError(f"source holds {nbytes} bytes, {count} frames need {count * self.fd.get_frame_size()}")
        _lib.call("hv_holo_store_frames", self.h, ctypes.c_void_p(ptr), count, _stream_arg(stream))

    def write_footer(self) -> None:
        """Compact JSON, keys sorted -- the form nlohmann::json::dump() gives."""
        s = dump_footer(self.meta_data)
        _lib.call("hv_holo_write_footer", self.h, s, len(s))

    def correct_number_of_frames(self, nb_frames_written: int) -> None:
        _lib.call("hv_holo_correct_number_of_frames", self.h, nb_frames_written)

    def close(self) -> None:
        if self.h:
            _lib.call("hv_holo_writer_close", self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dump_footer(meta: dict) -> bytes:
    return json.dumps(meta, sort_keys=True, separators=(",", ":"), ensure_ascii=False).encode("utf-8") if meta else b""


def _device_ptr(x):
    if hasattr(x, "data_ptr"):
        if not x.is_cuda or not x.is_contiguous():
            raise ValueError("expected a contiguous device tensor")
        return x.data_ptr(), x.numel() * x.element_size()
    return int(x), None


def _stream_arg(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


# -- compute settings <-> footer JSON ---------------------------------------------------------------

def _num(x):
    return int(x) if float(x).is_integer() else float(x)


def _f32(x: float) -> float:
    """A float setting as the reference stores it (C++ float, dumped as double)."""
    return float(np.float32(x))


def _view_window(contrast: dict | None = None) -> dict:
    # Backend/includes/struct/view_struct.hh ViewWindow / ViewContrast defaults
    c = {"enabled": True, "auto_refresh": True, "invert": False, "min": 1.0, "max": 65535.0}
    if contrast:
        c.update(contrast)
    return {"log_enabled": False, "contrast": c}


def _view_xyz(contrast: dict | None = None) -> dict:
    w = _view_window(contrast)
    w.update({"horizontal_flip": False, "rotation": 0.0, "output_image_accumulation": 1, "enabled": False})
    return w


def _cut_view(s: ComputeSettings) -> dict:
    w = _view_xyz()
    w["enabled"] = bool(s.cuts_enabled)
    return w


def _channel(shift: bool = False) -> dict:
    # Backend/includes/struct/composite_struct.hh CompositeChannel / CompositeH defaults
    c = {"frame_index": {"min": 0, "max": 0, "activated": False},
         "slider_threshold": {"min": 0.0, "max": 1.0},
         "threshold": {"min": _f32(0.001), "max": _f32(99.8)}}
    if shift:
        c["slider_shift"] = {"min": 0.0, "max": 1.0}
    return c


def compute_settings_to_json(s: ComputeSettings) -> dict:
    """hv_compute_settings as the reference's ComputeSettings JSON, version
    "v1" (Backend/includes/struct/compute_settings_struct.hh:22-40): every
    key of Rendering, Views, Composite and AdvancedSettings is present, so
    that the reference's from_json (which requires each key) accepts it;
    settings this path does not have keep the structs' defaults."""
    if s.time_transformation not in _TIME_WRITE or s.space_transformation not in _SPACE_WRITE:
        raise UnsupportedSettings("space/time transformation outside the path")
    if s.img_type not in _IMG_WRITE:
        raise UnsupportedSettings(f"image type {s.img_type} outside the path")
    auto = s.contrast_mode != HV_CONTRAST_MANUAL
    rendering = {
        "image_mode": "HOLOGRAM", "batch_size": int(s.batch_size),
        "time_transformation_stride": int(s.time_stride),
        "filter2d": {"enabled": False, "inner_radius": 0, "outer_radius": 1},
        "space_transformation": _SPACE_WRITE[s.space_transformation],
        "time_transformation": _TIME_WRITE[s.time_transformation],
        "time_transformation_size": int(s.time_transformation_size),
        "lambda": _f32(s.lambda_), "propagation_distance": _f32(s.z_distance),
        "convolution": {"type": "", "divide": False}, "input_filter": {"type": "None"},
    }
    xy = _view_xyz({"enabled": bool(s.contrast_enabled), "auto_refresh": auto, "invert": bool(s.contrast_invert),
                    "min": _f32(s.contrast_min), "max": _f32(s.contrast_max)})
    xy["log_enabled"] = bool(s.log_enabled)
    xy["output_image_accumulation"] = max(1, int(s.img_accu_level))
    view = {
        "image_type": _IMG_WRITE[s.img_type], "fft_shift": bool(s.fft_shift),
        "x": {"width": int(s.x_width), "start": int(s.x_start)},
        "y": {"width": int(s.y_width), "start": int(s.y_start)},
        "z": {"width": int(s.p_width), "start": int(s.p_start)}, "z2": {"width": 0, "start": 0},
        "window": {"xy": xy, "yz": _cut_view(s), "xz": _cut_view(s), "filter2d": _view_window()},
        "renorm": bool(s.renorm_enabled),
        "reticle": {"display_enabled": bool(s.contrast_reticle_enabled), "scale": _f32(s.contrast_reticle_scale)},
        "registration": {"registration_enabled": bool(s.registration_enabled),
                         "registration_zone": _f32(s.registration_zone)},
    }
    composite = {"mode": "RGB", "auto_weight": False,
                 "rgb": {"frame_index": {"min": 0, "max": 0}, "weight": {"r": 1.0, "g": 1.0, "b": 1.0}},
                 "hsv": {"h": _channel(True), "s": _channel(), "v": _channel()}}
    advanced = {"buffer_size": {"input": 512, "file": 512, "record": 1024, "output": 256,
                                "time_transformation_cuts": 512},
                "filter2d_smooth": {"low": 0, "high": 0},
                "contrast": {"lower": _f32(s.contrast_lower), "upper": _f32(s.contrast_upper),
                             "frame_index_offset": 2},
                "renorm_constant": int(s.renorm_constant), "raw_bitshift": 0, "nb_frames_to_record": 0}
    return {"version": "v1", "image_rendering": rendering, "view": view, "color_composite_image": composite,
            "advanced": advanced}


def _enum_name(value, names) -> str:
    """The JSON name as nlohmann's enum from_json reads it: a name of the
    table, else the table's first entry."""
    names = list(names)
    return value if isinstance(value, str) and value in names else names[0]


def compute_settings_from_json(cs: dict, fd: FrameDescriptor, info: dict | None = None,
                               base: ComputeSettings | None = None) -> tuple[ComputeSettings, list[str]]:
    """The footer's compute settings applied to hv_compute_settings (the
    reference's ComputeSettings::Update, compute_settings_struct.hh:42-47),
    with the frame geometry from the header and pixel size / camera fps from
    the footer info.  Settings the JSON does not carry (unwrap, arithmetic
    precisions, contrast subsample, the automatic contrast policy) come from
    `base` (default: hv_default_settings).  Returns (settings, ignored): `ignored` names the
    enabled steps the path does not run (band-pass filters, output
    flip/rotation); renormalisation, registration, log10, image accumulation
    and the cuts are mapped.
    Settings without version "v1" are upgraded first
    (footer_versions.upgrade_compute_settings; FooterUpgradeError where the
    reference's converter fails).  Enum names outside the reference's tables
    read as each table's first entry, as nlohmann does.
    Raises UnsupportedSettings for computations outside the path: raw mode,
    SSA_STFT, phase increase and composite images."""
    cs = upgrade_compute_settings(cs)  # v0 -> v1 (compute_settings.cc:75-77)
    if base is None:
        s = ComputeSettings()
        _lib.load().hv_default_settings(ctypes.byref(s))
    else:
        s = ComputeSettings.from_buffer_copy(base)
    s.width, s.height, s.depth = fd.width, fd.height, fd.depth
    ignored: list[str] = []
    r = cs.get("image_rendering", {})
    v = cs.get("view", {})
    adv = cs.get("advanced", {})
    # a name outside an enum's table reads as the table's first entry (nlohmann's enum from_json)
    mode = _enum_name(r.get("image_mode", "RAW"), ("RAW", "HOLOGRAM"))
    if mode != "HOLOGRAM":
        raise UnsupportedSettings(f"image_mode {mode}: only HOLOGRAM runs the compute path")
    space = _enum_name(r.get("space_transformation", "NONE"), _SPACE_NAMES)
    time = _enum_name(r.get("time_transformation", "NONE"), ("STFT", "PCA", "NONE", "SSA_STFT", "None"))
    if time not in _TIME_NAMES:
        raise UnsupportedSettings(f"time_transformation {time}")
    img = _enum_name(v.get("image_type", "MODULUS"), ("MODULUS", "SQUARED_MODULUS", "ARGUMENT", "PHASE_INCREASE",
                                                      "COMPOSITE", "MOMENTS_0", "MOMENTS_1", "MOMENTS_2"))
    if img not in _IMG_NAMES:
        raise UnsupportedSettings(f"image_type {img}")
    s.space_transformation = _SPACE_NAMES[space]
    s.time_transformation = _TIME_NAMES[time]
    s.img_type = _IMG_NAMES[img]
    s.batch_size = int(r.get("batch_size", s.batch_size))
    s.time_stride = int(r.get("time_transformation_stride", s.time_stride))
    s.time_transformation_size = int(r.get("time_transformation_size", s.time_transformation_size))
    s.lambda_ = float(r.get("lambda", s.lambda_))
    s.z_distance = float(r.get("propagation_distance", s.z_distance))
    z = v.get("z", {})
    s.p_start = int(z.get("start", s.p_start))
    s.p_width = int(z.get("width", s.p_width))
    s.fft_shift = int(bool(v.get("fft_shift", False)))
    xy = v.get("window", {}).get("xy", {})
    c = xy.get("contrast", {})
    s.contrast_enabled = int(bool(c.get("enabled", True)))
    s.contrast_invert = int(bool(c.get("invert", False)))
    s.contrast_min = float(c.get("min", s.contrast_min))
    s.contrast_max = float(c.get("max", s.contrast_max))
    # auto_refresh: the reference's percentile contrast, or the automatic mode `base` already selects
    if not c.get("auto_refresh", True):
        s.contrast_mode = HV_CONTRAST_MANUAL
    elif s.contrast_mode == HV_CONTRAST_MANUAL or base is None:
        s.contrast_mode = HV_CONTRAST_REFERENCE
    ac = adv.get("contrast", {})
    s.contrast_lower = float(ac.get("lower", s.contrast_lower))
    s.contrast_upper = float(ac.get("upper", s.contrast_upper))
    if info:
        if "pixel_size" in info:
            s.pixel_size = float(info["pixel_size"])
        if "camera_fps" in info:
            s.camera_fps = float(info["camera_fps"])
    # post-processing of the xy image (pipe.cc:353-370)
    s.renorm_enabled = int(bool(v.get("renorm", False)))
    s.renorm_constant = int(adv.get("renorm_constant", s.renorm_constant))
    s.log_enabled = int(bool(xy.get("log_enabled", False)))
    s.img_accu_level = max(1, int(xy.get("output_image_accumulation", 1) or 1))
    # XZ / YZ cuts at the x / y views
    win = v.get("window", {})
    s.cuts_enabled = int(bool(win.get("xz", {}).get("enabled") or win.get("yz", {}).get("enabled")))
    s.x_start, s.x_width = int(v.get("x", {}).get("start", 0)), int(v.get("x", {}).get("width", 0))
    s.y_start, s.y_width = int(v.get("y", {}).get("start", 0)), int(v.get("y", {}).get("width", 0))
    # steps outside the path that the footer switches on
    if xy.get("horizontal_flip") or float(xy.get("rotation", 0.0) or 0.0) != 0.0:
        ignored.append("view.window.xy.horizontal_flip/rotation")
    reg = v.get("registration", {})
    s.registration_enabled = int(bool(reg.get("registration_enabled", False)))
    s.registration_zone = float(reg.get("registration_zone", s.registration_zone))
    # the contrast reticle (Reticle::Load, all_struct.cc:221-225)
    ret = v.get("reticle", {})
    s.contrast_reticle_enabled = int(bool(ret.get("display_enabled", False)))
    s.contrast_reticle_scale = float(ret.get("scale", s.contrast_reticle_scale))
    if s.contrast_reticle_enabled and s.contrast_subsample > 1:
        s.contrast_subsample = 1
    if r.get("filter2d", {}).get("enabled"):
        ignored.append("image_rendering.filter2d")
    if r.get("input_filter", {}).get("type", "None") not in ("None", ""):
        ignored.append("image_rendering.input_filter")
    return s, ignored
