"""Upgrades of .holo footers written under header versions 2 to 4, and of
compute settings without a version, to the current layout.

Reference behaviour (HoloFileConverter::convert_holo_file,
Backend/sources/version/holo_file_converter.cc:54-124): versions below 2
count as 2; each version v < 7 is brought to v + 1 in turn -- 2 -> 3 and
3 -> 4 and 4 -> 5 by the JSON patches of
Backend/assets/json_patches_holofile/ (patch_v2_to_v3.json ...
patch_v4_to_v5.json, RFC 6902, applied by nlohmann::json::patch) plus the
converter functions around them (holo_file_converter.cc:21-52); 5 -> 6 and
6 -> 7 change nothing.  Any failing step makes the whole import fail and the
reference then loads no settings from the file (input_holo_file.cc:135-136).
Compute settings without "version": "v1" get the v0 -> v1 patch
(Backend/assets/json_patches_settings/patch_v1.json,
compute_settings_converter.cc:22-96) before they are merged into the
defaults.

The edit lists below state the same key moves as those patch files (the
format the files were written in fixes them), in order; `_apply` follows
nlohmann's patch semantics for the operations they use: "remove" and the
source of "move"/"copy" must exist, "add" needs the first path token to
exist and creates missing intermediate objects, "move" is remove + add.
"""
from __future__ import annotations

import copy

LATEST_VERSION = 7


class FooterUpgradeError(ValueError):
    """A footer or compute settings object that the version upgrade cannot
    convert (a key the upgrade moves or removes is missing, a value of the
    wrong type); the reference then imports no settings."""


# -- JSON pointer patch operations (RFC 6901 paths, nlohmann semantics) -------------------------

def _tokens(path: str) -> list[str]:
    if path == "":
        return []
    if not path.startswith("/"):
        raise FooterUpgradeError(f"invalid JSON pointer {path!r}")
    return [t.replace("~1", "/").replace("~0", "~") for t in path[1:].split("/")]


def _at(doc, toks: list[str]):
    cur = doc
    for t in toks:
        if isinstance(cur, dict) and t in cur:
            cur = cur[t]
        elif isinstance(cur, list) and t.isdigit() and int(t) < len(cur):
            cur = cur[int(t)]
        else:
            raise FooterUpgradeError(f"key '{t}' not found")
    return cur


def _remove(doc, path: str):
    toks = _tokens(path)
    if not toks:
        raise FooterUpgradeError("cannot remove the root")
    parent = _at(doc, toks[:-1])
    key = toks[-1]
    if isinstance(parent, dict):
        if key not in parent:
            raise FooterUpgradeError(f"key '{key}' not found")
        return parent.pop(key)
    if isinstance(parent, list) and key.isdigit() and int(key) < len(parent):
        return parent.pop(int(key))
    raise FooterUpgradeError(f"cannot remove {path!r}")


def _add(doc, path: str, value):
    toks = _tokens(path)
    if not toks:
        return value  # adding at the root replaces the document
    if len(toks) > 1:
        _at(doc, toks[:1])  # the first token must exist
    cur = doc
    for t in toks[:-1]:  # intermediate objects are created (json::operator[])
        if isinstance(cur, dict):
            if cur.get(t) is None:
                cur[t] = {}
            cur = cur[t]
        elif isinstance(cur, list) and t.isdigit() and int(t) < len(cur):
            cur = cur[int(t)]
        else:
            raise FooterUpgradeError(f"cannot add {path!r}")
    if not isinstance(cur, dict):
        raise FooterUpgradeError(f"cannot add {path!r}")
    cur[toks[-1]] = value
    return doc


def _apply(doc, edits):
    """edits: ("remove", path) | ("add", path, value) | ("move", from, to) | ("copy", from, to)."""
    for e in edits:
        if e[0] == "remove":
            _remove(doc, e[1])
        elif e[0] == "add":
            doc = _add(doc, e[1], copy.deepcopy(e[2]))
        elif e[0] == "move":
            v = _at(doc, _tokens(e[1]))
            _remove(doc, e[1])
            doc = _add(doc, e[2], v)
        elif e[0] == "copy":
            doc = _add(doc, e[2], copy.deepcopy(_at(doc, _tokens(e[1]))))
        else:
            raise AssertionError(e[0])
    return doc


# -- header version 2 -> 3: accumulation switches and the renormalisation constant dropped ------

_V2_TO_V3 = [("remove", "/" + k) for k in (
    "y_acc_enabled", "x_acc_enabled", "renorm_constant", "p_acc_enabled",
    "img_acc_slice_yz_enabled", "img_acc_slice_xz_enabled", "img_acc_slice_xy_enabled")]


# -- header version 3 -> 4: flat keys moved into an "info" and a "compute settings" tree ----------

def _v4_contrast():
    return {"auto refresh": True, "enabled": False, "invert": False, "max": 65535.0, "min": 1.0}


def _v4_view_window():
    return {"contrast": _v4_contrast(), "flip": False, "img accu level": 1, "log enabled": False, "rot": 0.0}


def _v4_thresholds():
    return {"slider threshold": {"max": 0.0, "min": 0.0}, "threshold": {"high": 0.0, "low": 0.0}}


def _v4_compute_settings() -> dict:
    """The defaults the 3 -> 4 upgrade adds before moving the file's values in."""
    sv = {"p": {"activated": False, "max": 0, "min": 0}, **_v4_thresholds()}
    return {
        "advanced": {
            "buffer size": {"file": 512, "input": 512, "output": 1024, "record": 256,
                            "time transformation cuts": 512},
            "contrast": {"cuts p offset": 2, "lower": 0.5, "upper": 99.5},
            "filter2d smooth": {"high": 0, "low": 0},
            "renorm constant": 5,
        },
        "composite": {
            "auto weight": False,
            "hsv": {"h": {"blur": {"enabled": False, "kernel size": 1}, "p": {"max": 0, "min": 0},
                          **_v4_thresholds()},
                    "s": copy.deepcopy(sv), "v": copy.deepcopy(sv)},
            "mode": "RGB",
            "rgb": {"p": {"max": 0, "min": 0}, "weight": {"b": 0.0, "g": 0.0, "r": 0.0}},
        },
        "image rendering": {
            "batch size": 1,
            "convolution": {"divide": False, "enabled": False, "type": "None"},
            "filter2d": {"enabled": False, "n1": 0, "n2": 1},
            "image mode": "RAW",
            "lambda": 8.520000278622319e-07,
            "space transformation": "NONE",
            "time transformation": "NONE",
            "time transformation size": 1,
            "time transformation stride": 1,
            "z distance": 1.5,
        },
        "version": "v1",
        "view": {
            "fft shift": False,
            "p": {"accu level": 0, "index": 0},
            "q": {"accu level": 0, "index": 0},
            "renorm": True,
            "reticle": {"display enabled": False, "scale": 0.5},
            "type": "MODULUS",
            "window": {"filter2d": {"contrast": _v4_contrast(), "log enabled": False},
                       "xy": _v4_view_window(), "xz": _v4_view_window(), "yz": _v4_view_window()},
            "x": {"accu level": 0, "cuts": 0},
            "y": {"accu level": 0, "cuts": 0},
        },
    }


_CS4 = "/compute settings"
_V3_TO_V4 = [
    ("add", "/info", {"contiguous": 1, "input fps": 1, "pixel size": {"x": 1, "y": 1}, "raw bitshift": 0}),
    ("add", _CS4, None),  # value: _v4_compute_settings()
] + [("move", "/" + src, _CS4 + dst) for src, dst in (
    ("#img", "/image rendering/time transformation size"),
    ("algorithm", "/image rendering/space transformation"),
    ("mode", "/image rendering/image mode"),
    ("time_filter", "/image rendering/time transformation"),
    ("contrast_max", "/view/window/xy/contrast/max"),
    ("contrast_min", "/view/window/xy/contrast/min"),
    ("fft_shift_enabled", "/view/fft shift"),
    ("img_acc_slice_xy_level", "/view/window/xy/img accu level"),
    # sic: the reference's patch writes this level under a key without spaces, so the
    # yz level keeps its default and the file's value stays beside it
    ("img_acc_slice_yz_level", "/view/window/yz/imgacculevel"),
    ("img_acc_slice_xz_level", "/view/window/xz/img accu level"),
    ("lambda", "/image rendering/lambda"),
    ("log_scale", "/view/window/xy/log enabled"),
    ("p", "/view/p/index"),
    ("p_acc_level", "/view/p/accu level"),
    ("renorm_enabled", "/view/renorm"),
    ("x_acc_level", "/view/x/accu level"),
    ("y_acc_level", "/view/y/accu level"),
    ("z", "/image rendering/z distance"),
)] + [
    ("copy", "/pixel_size", "/info/pixel size/y"),
    ("move", "/pixel_size", "/info/pixel size/x"),
]

# enum values as nlohmann serialises them (the first name of a value; a value
# outside the table takes the table's first name): SpaceTransformation
# (enum_space_transformation.hh:31-42), Computation (enum_computation.hh:23-26),
# TimeTransformation (enum_time_transformation.hh:25-31)
_SPACE_JSON = ("NONE", {0: "NONE", 1: "FRESNELTR", 2: "ANGULARSP"})
_COMPUTATION_JSON = ("RAW", {0: "RAW", 1: "HOLOGRAM"})
_TIME_JSON = ("STFT", {0: "NONE", 1: "STFT", 2: "PCA", 3: "SSA_STFT"})


def _json_int(v) -> int:
    """static_cast<int>(json): numbers (floats truncated) and booleans."""
    if isinstance(v, bool):
        return int(v)
    if isinstance(v, (int, float)):
        return int(v)
    raise FooterUpgradeError(f"type must be number, but is {type(v).__name__}")


def _enum_json(table, v: int) -> str:
    return table[1].get(v, table[0])


def _v3_to_v4(meta):
    edits = list(_V3_TO_V4)
    edits[1] = ("add", _CS4, _v4_compute_settings())
    meta = _apply(meta, edits)
    r = meta["compute settings"]["image rendering"]
    r["space transformation"] = _enum_json(_SPACE_JSON, _json_int(r["space transformation"]))
    r["image mode"] = _enum_json(_COMPUTATION_JSON, _json_int(r["image mode"]) - 1)
    r["time transformation"] = _enum_json(_TIME_JSON, _json_int(r["time transformation"]))
    return meta


# -- header version 4 -> 5: keys with spaces renamed, "compute settings" -> "compute_settings" -----

_CS5 = "compute_settings"
# (parent object, old key, new key), in order: later entries use earlier new names
_V4_TO_V5_RENAMES = [
    ("", "compute settings", _CS5),
    (f"{_CS5}/advanced", "buffer size", "buffer_size"),
    (f"{_CS5}/advanced", "filter2d smooth", "filter2d_smooth"),
    (f"{_CS5}/advanced", "renorm constant", "renorm_constant"),
    (f"{_CS5}/advanced/contrast", "cuts p offset", "frame_index_offset"),
    (f"{_CS5}/advanced/buffer_size", "time transformation cuts", "time_transformation_cuts"),
    (_CS5, "composite", "color_composite_image"),
]
_CC = f"{_CS5}/color_composite_image"
_V4_TO_V5_RENAMES += [
    (_CC, "auto weight", "auto_weight"),
    (f"{_CC}/hsv/h/blur", "kernel size", "kernel_size"),
] + [(f"{_CC}/hsv/{c}", "slider threshold", "slider_threshold") for c in "hsv"] \
  + [(f"{_CC}/hsv/{c}/threshold", old, new) for c in "hsv" for old, new in (("high", "max"), ("low", "min"))] \
  + [(f"{_CC}/hsv/{c}", "p", "frame_index") for c in "hsv"] \
  + [(f"{_CC}/rgb", "p", "frame_index"), (_CS5, "image rendering", "image_rendering")]
_IR = f"{_CS5}/image_rendering"
_V4_TO_V5_RENAMES += [(_IR, old, new) for old, new in (
    ("batch size", "batch_size"), ("image mode", "image_mode"), ("space transformation", "space_transformation"),
    ("time transformation", "time_transformation"), ("time transformation size", "time_transformation_size"),
    ("time transformation stride", "time_transformation_stride"), ("z distance", "propagation_distance"))] + [
    (f"{_IR}/filter2d", "n1", "inner_radius"), (f"{_IR}/filter2d", "n2", "outer_radius")]
_VW = f"{_CS5}/view"
_WIN = f"{_VW}/window"
_V4_TO_V5_RENAMES += [
    (_VW, "fft shift", "fft_shift"),
    (_VW, "p", "z"), (f"{_VW}/z", "accu level", "width"), (f"{_VW}/z", "index", "start"),
    (_VW, "q", "z2"), (f"{_VW}/z2", "accu level", "width"), (f"{_VW}/z2", "index", "start"),
    (f"{_VW}/reticle", "display enabled", "display_enabled"),
    (f"{_WIN}/filter2d/contrast", "auto refresh", "auto_refresh"),
    (f"{_WIN}/filter2d", "log enabled", "log_enabled"),
    (f"{_WIN}/xy", "log enabled", "log_enabled"),
    (f"{_WIN}/xz", "log enabled", "log_enabled"),
    (f"{_WIN}/yz", "log enabled", "log_enabled"),
    (f"{_WIN}/xy", "img accu level", "output_image_accumulation"),
    (f"{_WIN}/xz", "img accu level", "output_image_accumulation"),
    (f"{_WIN}/yz/contrast", "auto refresh", "auto_refresh"),
    (f"{_WIN}/xy/contrast", "auto refresh", "auto_refresh"),
    (f"{_WIN}/xz/contrast", "auto refresh", "auto_refresh"),
    (f"{_WIN}/yz", "img accu level", "output_image_accumulation"),
    (f"{_VW}/x", "accu level", "width"), (f"{_VW}/y", "accu level", "width"),
    (f"{_VW}/x", "cuts", "start"), (f"{_VW}/y", "cuts", "start"),
    (_VW, "type", "image_type"),
] + [(f"{_WIN}/{w}", "flip", "horizontal_flip") for w in ("xy", "xz", "yz")] \
  + [(f"{_WIN}/{w}", "rot", "rotation") for w in ("xy", "xz", "yz")] \
  + [("info", "input fps", "input_fps"), ("info", "pixel size", "pixel_pitch")]


def _path(parent: str, key: str) -> str:
    return "/" + (parent + "/" if parent else "") + key


_V4_TO_V5 = [("move", _path(p, old), _path(p, new)) for p, old, new in _V4_TO_V5_RENAMES] + [
    ("move", "/info/raw bitshift", f"/{_CS5}/advanced/raw_bitshift")]


def _v4_to_v5(meta):
    if isinstance(meta, dict) and "file info" in meta:
        info = copy.deepcopy(meta["file info"])
        if info is None:
            info = {}
        if not isinstance(info, dict):
            raise FooterUpgradeError("\"file info\" is not an object")
        info["input fps"] = 1
        info["contiguous"] = 1
        meta["info"] = info
    meta = _apply(meta, _V4_TO_V5)
    # sic: the reference renames the old image type names under "compute_setting"
    # (holo_file_converter.cc:46-49), a key its own patch never writes; the lookup
    # leaves that key behind with a null image type and renames nothing
    cs = meta.get("compute_setting")
    if cs is None:
        cs = meta["compute_setting"] = {}
    if not isinstance(cs, dict):
        raise FooterUpgradeError("\"compute_setting\" is not an object")
    view = cs.get("view")
    if view is None:
        view = cs["view"] = {}
    if not isinstance(view, dict):
        raise FooterUpgradeError("\"compute_setting/view\" is not an object")
    t = view.setdefault("image_type", None)
    if t == "PHASEINCREASE":
        view["image_type"] = "PHASE_INCREASE"
    elif t == "SQUAREDMODULUS":
        view["image_type"] = "SQUARED_MODULUS"
    return meta


_STEPS = {2: lambda m: _apply(m, _V2_TO_V3), 3: _v3_to_v4, 4: _v4_to_v5, 5: lambda m: m, 6: lambda m: m}


def upgrade_footer(meta, version: int):
    """The footer `meta` (parsed JSON) of a header of `version`, brought to
    the version-7 layout (a new object; `meta` is not changed).  Raises
    FooterUpgradeError where the reference's converter fails (versions above
    7 included)."""
    if version > LATEST_VERSION:
        raise FooterUpgradeError(f"holo file version {version} not supported, latest supported is {LATEST_VERSION}")
    meta = copy.deepcopy(meta)
    for v in range(max(version, 2), LATEST_VERSION):
        if not isinstance(meta, (dict, list)) and v <= 4:
            raise FooterUpgradeError("footer is not a JSON object")
        meta = _STEPS[v](meta)
    return meta


# -- compute settings v0 -> v1 (json_patches_settings/patch_v1.json) -----------------------------

_SETTINGS_V0_TO_V1 = [
    ("remove", "/image_rendering/input_filter/enabled"),
    ("remove", "/image_rendering/convolution/enabled"),
    ("remove", "/advanced/record_mode"),
    ("add", "/version", "v1"),
    ("add", "/view/registration", {"registration_enabled": False, "registration_zone": 0.7}),
]


def upgrade_compute_settings(cs: dict) -> dict:
    """Compute settings brought to version "v1" (a new object): any
    "version" other than "v1" (none included) reads as v0 and gets the
    v0 -> v1 edits.  Raises FooterUpgradeError where the reference's
    ComputeSettingsConverter fails."""
    if isinstance(cs, dict) and cs.get("version") == "v1":
        return cs
    return _apply(copy.deepcopy(cs), _SETTINGS_V0_TO_V1)
