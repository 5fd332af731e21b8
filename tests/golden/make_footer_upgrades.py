"""Writes footer_upgrades.json: footers of .holo header versions 1 to 6 and
compute settings without a version, each with the document the reference's
upgrade produces (or null where it fails).  Run in a tree that has the
reference at /root/reference; the output is committed.

The expected documents come from the reference's own JSON patch files
(Backend/assets/json_patches_holofile/patch_v2_to_v3.json,
patch_v3_to_v4.json, patch_v4_to_v5.json and
Backend/assets/json_patches_settings/patch_v1.json), read as data and applied
here with a separate RFC 6902 evaluator, around the steps of the converter
functions (Backend/sources/version/holo_file_converter.cc:21-52,
compute_settings_converter.cc:22-96).  The inputs are synthetic: the
reference ships no recording of those versions outside Git LFS.
"""
import copy
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ASSETS = "/root/reference/Backend/assets"


class Fail(Exception):
    pass


def unescape(tok):
    return tok.replace("~1", "/").replace("~0", "~")


def pointer(p):
    return [] if p == "" else [unescape(t) for t in p.split("/")[1:]]


def resolve(doc, toks):
    for t in toks:
        if not isinstance(doc, dict) or t not in doc:
            raise Fail(t)
        doc = doc[t]
    return doc


def rfc6902(doc, patch):
    """nlohmann::json::patch for objects: remove/move/copy need their source,
    add needs the first token and creates the objects between."""
    for op in patch:
        toks = pointer(op["path"])
        if op["op"] in ("move", "copy"):
            val = copy.deepcopy(resolve(doc, pointer(op["from"])))
            if op["op"] == "move":
                src = pointer(op["from"])
                del resolve(doc, src[:-1])[src[-1]]
        elif op["op"] == "remove":
            parent = resolve(doc, toks[:-1])
            if toks[-1] not in parent:
                raise Fail(toks[-1])
            del parent[toks[-1]]
            continue
        elif op["op"] == "add":
            val = copy.deepcopy(op["value"])
        else:
            raise NotImplementedError(op["op"])
        if not toks:
            doc = val
            continue
        if len(toks) > 1:
            resolve(doc, toks[:1])
        node = doc
        for t in toks[:-1]:
            if node.get(t) is None:
                node[t] = {}
            node = node[t]
        node[toks[-1]] = val
    return doc


def as_int(v):
    if isinstance(v, (bool, int, float)):
        return int(v)
    raise Fail("not a number")


def patch_file(name):
    with open(os.path.join(ASSETS, name)) as f:
        return json.load(f)


def upgrade(meta, version):
    if version > 7:
        return None
    meta = copy.deepcopy(meta)
    try:
        for v in range(max(2, version), 7):
            if v == 2:
                meta = rfc6902(meta, patch_file("json_patches_holofile/patch_v2_to_v3.json"))
            elif v == 3:
                meta = rfc6902(meta, patch_file("json_patches_holofile/patch_v3_to_v4.json"))
                ir = meta["compute settings"]["image rendering"]
                space = as_int(ir["space transformation"])
                ir["space transformation"] = ["NONE", "FRESNELTR", "ANGULARSP"][space] if 0 <= space <= 2 else "NONE"
                mode = as_int(ir["image mode"]) - 1
                ir["image mode"] = ["RAW", "HOLOGRAM"][mode] if 0 <= mode <= 1 else "RAW"
                tt = as_int(ir["time transformation"])
                ir["time transformation"] = ["NONE", "STFT", "PCA", "SSA_STFT"][tt] if 0 <= tt <= 3 else "STFT"
            elif v == 4:
                if "file info" in meta:
                    info = copy.deepcopy(meta["file info"]) or {}
                    info["input fps"] = 1
                    info["contiguous"] = 1
                    meta["info"] = info
                meta = rfc6902(meta, patch_file("json_patches_holofile/patch_v4_to_v5.json"))
                view = meta.setdefault("compute_setting", {}).setdefault("view", {})
                view.setdefault("image_type", None)
                if view["image_type"] == "PHASEINCREASE":
                    view["image_type"] = "PHASE_INCREASE"
                elif view["image_type"] == "SQUAREDMODULUS":
                    view["image_type"] = "SQUARED_MODULUS"
    except (Fail, KeyError, TypeError, AttributeError):
        return None
    return meta


def upgrade_settings(cs):
    if cs.get("version") == "v1":
        return cs
    try:
        return rfc6902(copy.deepcopy(cs), patch_file("json_patches_settings/patch_v1.json"))
    except (Fail, KeyError, TypeError, AttributeError):
        return None


def v3_footer(**over):
    d = {"#img": 16, "algorithm": 2, "mode": 2, "time_filter": 1, "contrast_max": 3.5, "contrast_min": 0.25,
         "fft_shift_enabled": True, "img_acc_slice_xy_level": 4, "img_acc_slice_yz_level": 3,
         "img_acc_slice_xz_level": 2, "lambda": 7.85e-07, "log_scale": False, "p": 3, "p_acc_level": 5,
         "renorm_enabled": False, "x_acc_level": 1, "y_acc_level": 2, "z": 0.25, "pixel_size": 20.0}
    d.update(over)
    return {k: v for k, v in d.items() if v is not None}


def v2_footer(**over):
    d = v3_footer(**over)
    d.update({"y_acc_enabled": False, "x_acc_enabled": False, "renorm_constant": 5, "p_acc_enabled": True,
              "img_acc_slice_yz_enabled": False, "img_acc_slice_xz_enabled": False,
              "img_acc_slice_xy_enabled": True})
    return d


def v4_footer(info_key="info", image_type="MODULUS"):
    """A version-4 document: the 3 -> 4 patch applied to a version-3 footer, then edited."""
    d = rfc6902(v3_footer(), patch_file("json_patches_holofile/patch_v3_to_v4.json"))
    ir = d["compute settings"]["image rendering"]
    ir.update({"space transformation": "FRESNELTR", "image mode": "HOLOGRAM", "time transformation": "PCA",
               "batch size": 8, "time transformation stride": 8})
    d["compute settings"]["view"]["type"] = image_type
    d["compute settings"]["view"]["p"] = {"accu level": 2, "index": 5}
    d["info"]["input fps"] = 250
    d["info"]["pixel size"] = {"x": 12.0, "y": 12.0}
    if info_key != "info":
        d[info_key] = d.pop("info")
    return d


def main():
    v5 = {"compute_settings": {"version": "v1", "image_rendering": {"batch_size": 4}},
          "info": {"input_fps": 30, "pixel_pitch": {"x": 5.5, "y": 5.5}}}
    footers = [
        ("v2 complete", 2, v2_footer()),
        ("v1 header counts as v2", 1, v2_footer()),
        ("v2 without a v2 key", 2, v3_footer()),
        ("v3 complete", 3, v3_footer()),
        ("v3 Fresnel PCA raw", 3, v3_footer(algorithm=1, mode=1, time_filter=2)),
        ("v3 enums outside the tables", 3, v3_footer(algorithm=7, mode=0, time_filter=9)),
        ("v3 float and boolean enums", 3, v3_footer(algorithm=2.9, mode=True, time_filter=False)),
        ("v3 string enum", 3, v3_footer(algorithm="FFT2")),
        ("v3 missing z", 3, v3_footer(z=None)),
        ("v3 existing info replaced", 3, dict(v3_footer(), info={"camera": "x"})),
        ("v4 complete", 4, v4_footer()),
        ("v4 file info", 4, v4_footer(info_key="file info")),
        ("v4 squared modulus name", 4, v4_footer(image_type="SQUAREDMODULUS")),
        ("v4 with compute_setting", 4, dict(v4_footer(), compute_setting={"view": {"image_type": "PHASEINCREASE"}})),
        ("v4 in v5 layout", 4, v5),
        ("v5", 5, v5),
        ("v6", 6, v5),
        ("v8", 8, v5),
        ("v3 not an object", 3, [1, 2]),
    ]
    v0 = {"image_rendering": {"input_filter": {"enabled": False, "type": "None"},
                              "convolution": {"enabled": False, "type": "None", "divide": False}},
          "advanced": {"record_mode": "RAW", "renorm_constant": 5}, "view": {"fft_shift": True}}
    settings = [
        ("v0 complete", v0),
        ("v0 named", dict(v0, version="v0")),
        ("unknown version reads as v0", dict(v0, version="v9")),
        ("v0 missing record_mode", {**v0, "advanced": {"renorm_constant": 5}}),
        ("v1", {"version": "v1", "view": {"fft_shift": True}}),
    ]
    out = {"footers": [{"case": c, "version": v, "input": m, "expected": upgrade(m, v)} for c, v, m in footers],
           "settings": [{"case": c, "input": m, "expected": upgrade_settings(m)} for c, m in settings]}
    with open(os.path.join(HERE, "footer_upgrades.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    for e in out["footers"] + out["settings"]:
        print(e["case"], "->", "fails" if e["expected"] is None else "ok")


if __name__ == "__main__":
    main()
