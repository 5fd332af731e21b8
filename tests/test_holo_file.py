""".holo reader/writer on the CPU: the native reader and writer (hv_holo_*)
against the format restatement in oracle/holo_format.py, the oracle against
the reference's own description of the header (its Python test harness) and
the sizes of its test recordings, and the footer's compute settings against
the reference's CLI test settings files.  Host I/O only: no GPU call."""
import ctypes
import glob
import json
import os
import re
import struct

import numpy as np
import pytest

from oracle import holo_format as hf

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def holo():
    from holovibes_amd import holo_file
    holo_file._lib.load()
    return holo_file


def _frames(n, h, w, depth, seed=0):
    rng = np.random.default_rng(seed)
    dt = {1: np.uint8, 2: np.uint16, 4: np.float32}[depth]
    if depth == 4:
        return rng.standard_normal((n, h, w)).astype(dt)
    return rng.integers(0, np.iinfo(dt).max + 1, (n, h, w), dtype=dt)


# ---- the oracle, pinned -------------------------------------------------------------------------

def test_oracle_header_matches_reference_harness_layout():
    """The reference's Python harness reads the same header as
    '=4sHHIIIQB' followed by padding to 64 bytes (tests/holo.py:18-33): the
    first 29 bytes of the oracle's header unpack to the same fields there."""
    harness = "=4sHHIIIQB"
    assert struct.calcsize(harness) + 35 == 64
    h = hf.pack_header(48, 20, 2, 7, hf.PROCESSED)
    magic, version, bpp, w, hh, n, total, endian = struct.unpack(harness, h[:29])
    assert (magic, version, bpp, w, hh, n, total, endian) == (b"HOLO", 7, 16, 48, 20, 7, 48 * 20 * 2 * 7, 0)
    assert h[29] == hf.PROCESSED and h[30:] == b"\0" * 34


def test_oracle_matches_reference_recording_sizes():
    """Each recording the reference's tests read ("<n>frame(s)_<W>_<H>.holo")
    is 64 + n*W*H*depth bytes for one pixel depth of 1 or 2 bytes: a 64-byte
    header, frames back to back, no footer."""
    recs = json.load(open(os.path.join(GOLDEN, "holo_recordings.json")))
    checked = 0
    for r in recs:
        m = re.match(r"(\d+)frames?_(\d+)_(\d+)\.holo$", r["name"])
        if not m:
            continue
        n, w, h = map(int, m.groups())
        depths = [d for d in (1, 2) if len(hf.pack_header(w, h, d, n)) + n * w * h * d == r["size"]]
        assert len(depths) == 1, r
        data = hf.holo_bytes(np.zeros((n, h, w), {1: np.uint8, 2: np.uint16}[depths[0]]), depths[0])
        assert len(data) == r["size"]
        assert hf.parse(data)["footer"] == b""
        checked += 1
    assert checked >= 5


def test_oracle_checks():
    data = hf.holo_bytes(_frames(3, 8, 16, 2), 2, footer={"a": 1})
    p = hf.parse(data)
    assert (p["img_nb"], p["frames_in_file"], p["footer"]) == (3, 3, b'{"a":1}')
    with pytest.raises(hf.OracleFormatError, match="bytecode size"):
        hf.parse(b"HOLA" + data[4:])
    with pytest.raises(hf.OracleFormatError, match="total data size"):
        hf.parse(data[:20] + struct.pack("<Q", 5) + data[28:])


# ---- native reader / writer against the oracle -------------------------------------------------

def test_make_header_matches_oracle(holo):
    from holovibes_amd._lib import HoloHeader, call
    assert ctypes.sizeof(HoloHeader) == 64
    assert HoloHeader.total_data_size.offset == 20 and HoloHeader.endianness.offset == 28
    for (w, h, d, n, t) in [(64, 32, 2, 3, 1), (2048, 2048, 2, 512, 0), (512, 512, 4, 384, 2), (1, 1, 1, 0, 1)]:
        hdr = HoloHeader()
        call("hv_holo_make_header", w, h, d, n, t, ctypes.byref(hdr))
        assert bytes(hdr) == hf.pack_header(w, h, d, n, t)


@pytest.mark.parametrize("depth,dtype", [(1, hf.RAW), (2, hf.PROCESSED), (4, hf.MOMENTS)])
def test_writer_bytes_match_oracle(holo, tmp_path, depth, dtype):
    frames = _frames(5, 24, 40, depth, seed=depth)
    fd = holo.FrameDescriptor(40, 24, depth)
    path = str(tmp_path / "w.holo")
    with holo.OutputHoloFile(path, fd, 5, dtype) as out:
        out.meta_data = {"info": {"input_fps": 10}, "compute_settings": {"z": [1, 2.5]}}
        out.write_header()
        out.write_frame(frames[:2])
        out.write_frame(frames[2:])
        out.write_footer()
    ref = hf.holo_bytes(frames, depth, dtype, footer=holo.dump_footer(out.meta_data))
    assert open(path, "rb").read() == ref


@pytest.mark.parametrize("n,h,w,depth", [(4, 16, 32, 2), (1, 32, 32, 1), (0, 8, 8, 2), (7, 33, 17, 4)])
def test_reader_matches_oracle(holo, tmp_path, n, h, w, depth):
    frames = _frames(n, h, w, depth, seed=n)
    data = hf.holo_bytes(frames, depth, hf.PROCESSED, footer={"compute_settings": {}, "info": {"input_fps": 3}})
    path = str(tmp_path / "r.holo")
    open(path, "wb").write(data)
    p = hf.parse(data)
    with holo.InputHoloFile(path) as f:
        i = f.info
        assert (i.width, i.height, i.depth, i.img_nb, i.version, i.bits_per_pixel, i.data_type) == \
            (p["width"], p["height"], p["depth"], p["img_nb"], p["version"], p["bits_per_pixel"], p["data_type"])
        assert (i.frame_size, i.total_data_size, i.frames_in_file, i.footer_size) == \
            (p["frame_size"], p["total_data_size"], p["frames_in_file"], len(p["footer"]))
        got = f.read_frames(n + 3)
        assert got.shape[0] == n and got.tobytes() == p["frames"]
        assert f.footer_bytes() == p["footer"]
        assert f.import_compute_settings() == {}
        # no pixel_pitch in the info: the reference's import_info sets nothing
        assert f.import_info() == {}


def test_reader_positions(holo, tmp_path):
    frames = _frames(10, 8, 8, 2)
    path = str(tmp_path / "p.holo")
    open(path, "wb").write(hf.holo_bytes(frames, 2))
    with holo.InputHoloFile(path) as f:
        f.set_pos_to_frame(3)
        a = f.read_frames(4)
        b = f.read_frames(10)
        assert np.array_equal(a, frames[3:7]) and np.array_equal(b, frames[7:])
        assert f.read_frames(2).shape[0] == 0
        with pytest.raises(holo.HoloFormatError, match="seek"):
            f.set_pos_to_frame(11)


def test_reader_refuses_like_reference(holo, tmp_path):
    good = hf.holo_bytes(_frames(2, 8, 8, 2), 2)
    cases = {
        "magic": (b"HOLA" + good[4:], "invalid bytecode size"),
        "short": (good[:40], "invalid bytecode size"),
        "total": (good[:20] + struct.pack("<Q", 7) + good[28:], "invalid total data size"),
    }
    for name, (data, msg) in cases.items():
        # the oracle agrees with the native reader on every refusal
        with pytest.raises(hf.OracleFormatError, match=msg):
            hf.parse(data)
        path = str(tmp_path / f"{name}.holo")
        open(path, "wb").write(data)
        with pytest.raises(holo.HoloFormatError, match=msg):
            holo.InputHoloFile(path)
    with pytest.raises(Exception, match="HV_ERR_IO"):
        holo.InputHoloFile(str(tmp_path / "missing.holo"))


def test_truncated_recording(holo, tmp_path):
    """A file cut inside its frames keeps its whole frames; its footer offset
    lies past the end, which import refuses ("invalid footer")."""
    frames = _frames(6, 8, 8, 2)
    data = hf.holo_bytes(frames, 2)[: 64 + 3 * 128 + 50]
    path = str(tmp_path / "t.holo")
    open(path, "wb").write(data)
    p = hf.parse(data)
    assert p["truncated"] and p["frames_in_file"] == 3
    with holo.InputHoloFile(path) as f:
        assert f.info.truncated == 1 and f.info.frames_in_file == 3 and f.get_total_nb_frames() == 6
        assert np.array_equal(f.read_frames(6), frames[:3])
        with pytest.raises(holo.HoloFormatError, match="invalid footer"):
            f.import_compute_settings()


def test_footer_versions_and_parse_failures(holo, tmp_path):
    frames = _frames(1, 8, 8, 1)
    meta = {"compute_settings": {"version": "v1"}, "info": {"pixel_pitch": {"x": 5.5, "y": 5.5}, "input_fps": 7}}

    def write(version, footer):
        path = str(tmp_path / f"v{version}.holo")
        open(path, "wb").write(hf.holo_bytes(frames, 1, hf.PROCESSED, footer=footer, version=version))
        return holo.InputHoloFile(path)

    with write(7, meta) as f:
        assert f.data_type == hf.PROCESSED
        assert f.import_compute_settings() == {"version": "v1"}
        assert f.import_info() == {"pixel_size": 5.5, "input_fps": 7, "camera_fps": 7}
    with write(6, meta) as f:  # data type in the header from version 7 only
        assert f.data_type == hf.RAW and f.import_compute_settings() == {"version": "v1"}
    with write(8, meta) as f:  # newer than the converter knows
        assert f.import_compute_settings() == {}
    with write(4, meta) as f:  # a version-7 layout under a version-4 header: the upgrade fails
        assert f.import_compute_settings() == {}
    with write(7, b"{not json") as f:  # skipped, as the reference does
        assert f.has_footer and f.import_compute_settings() == {} and f.import_info() == {}
    with write(7, None) as f:
        assert not f.has_footer and f.import_compute_settings() == {}


def test_correct_number_of_frames(holo, tmp_path):
    frames = _frames(3, 8, 8, 2)
    path = str(tmp_path / "c.holo")
    fd = holo.FrameDescriptor(8, 8, 2)
    with holo.OutputHoloFile(path, fd, 10) as out:
        out.meta_data = {"info": {"input_fps": 1}}
        out.write_header()
        out.write_frame(frames)
        out.write_footer()
        out.correct_number_of_frames(3)
    data = open(path, "rb").read()
    assert data == hf.holo_bytes(frames, 2, footer=holo.dump_footer({"info": {"input_fps": 1}}))
    with holo.InputHoloFile(path) as f:
        assert f.get_total_nb_frames() == 3 and f.meta_data == {}
        f.import_compute_settings()
        assert f.meta_data == {"info": {"input_fps": 1}}


# ---- footer compute settings ---------------------------------------------------------------------

_SPACE_BY_NAME = {"none": 0, "1fft": 1, "2fft": 2}
_TIME_BY_NAME = {"none": 0, "stft": 1, "pca": 2}


def _settings_files():
    return sorted(glob.glob(os.path.join(GOLDEN, "holo_settings", "*.json")))


def test_settings_files_map_as_their_names_say(holo):
    """The reference's CLI test cases are named for what they set
    ("space_1fft_time_stft", "time_stride_3", "batch_size_30", "raw"): the
    mapping gives those settings, and refuses what the path does not run."""
    fd = holo.FrameDescriptor(64, 64, 2)
    seen = 0
    for path in _settings_files():
        name = os.path.basename(path)[:-5]
        cs = json.load(open(path))
        m = re.search(r"space_(none|1fft|2fft)_time_(none|stft|pca|ssatft)", name)
        if "_raw" in name:
            with pytest.raises(holo.UnsupportedSettings, match="RAW"):
                holo.compute_settings_from_json(cs, fd)
            seen += 1
            continue
        if m and m.group(2) == "ssatft":
            with pytest.raises(holo.UnsupportedSettings, match="SSA_STFT"):
                holo.compute_settings_from_json(cs, fd)
            seen += 1
            continue
        s, ignored = holo.compute_settings_from_json(cs, fd)
        assert (s.width, s.height, s.depth) == (64, 64, 2)
        if m:
            assert s.space_transformation == _SPACE_BY_NAME[m.group(1)], name
            assert s.time_transformation == _TIME_BY_NAME[m.group(2)], name
        if "time_stride_3" in name:
            assert s.time_stride == 3
        if "batch_size_30" in name:
            assert s.batch_size == 30
        # the files set renormalisation ("renorm") and its constant explicitly
        assert s.renorm_enabled == int(cs["view"]["renorm"])
        assert s.renorm_constant == cs["advanced"]["renorm_constant"]
        assert ignored == []
        seen += 1
    assert seen == 16


def _key_tree(x):
    if isinstance(x, dict):
        return {k: _key_tree(v) for k, v in x.items()}
    return type(x).__name__ if not isinstance(x, (int, float)) or isinstance(x, bool) else "number"


def test_export_has_reference_key_tree(holo):
    """Every key of the reference's settings files is written, with the same
    kind of value, and no other: the reference's from_json requires each."""
    from holovibes_amd.configs import settings_for
    ours = _key_tree(holo.compute_settings_to_json(settings_for("c1")))
    for path in _settings_files():
        assert _key_tree(json.load(open(path))) == ours, path


@pytest.mark.parametrize("name", ["c1", "metric", "c2", "c3", "c5_stft", "c5_pca"])
def test_settings_round_trip(holo, name):
    from holovibes_amd.configs import settings_for
    s = settings_for(name)
    js = json.loads(json.dumps(holo.compute_settings_to_json(s)))
    fd = holo.FrameDescriptor(s.width, s.height, s.depth)
    back, ignored = holo.compute_settings_from_json(js, fd, {"pixel_size": s.pixel_size,
                                                              "camera_fps": s.camera_fps}, base=s)
    assert ignored == []
    for f, _ in type(s)._fields_:
        assert getattr(back, f) == getattr(s, f), f


def test_footer_dump_is_compact_sorted(holo, tmp_path):
    from holovibes_amd.configs import settings_for
    s = settings_for("c1")
    out = holo.OutputHoloFile(str(tmp_path / "f.holo"), holo.FrameDescriptor(256, 256, 2), 0)
    out.export_compute_settings(s, input_fps=1000, contiguous=16)
    raw = holo.dump_footer(out.meta_data)
    out.close()
    assert b'": ' not in raw and b', "' not in raw and b"\n" not in raw
    meta = json.loads(raw)
    assert list(meta) == sorted(meta) and meta["info"]["contiguous"] == 16
    assert meta["info"]["input_fps"] == 1000 and meta["info"]["camera_fps"] == 1000
    assert meta["compute_settings"]["image_rendering"]["lambda"] == float(np.float32(852e-9))


def test_large_writes_through_mapping(holo, tmp_path):
    """A write of 256 MiB or more reserves the file's blocks and goes
    through a shared mapping, smaller ones through pwrite; the file is byte
    for byte the one the oracle writes, footer and frame-count correction
    included."""
    rng = np.random.default_rng(11)
    frames = rng.integers(0, 65536, (132, 1024, 1024), dtype=np.uint16)  # 129 frames = 258 MiB in one call
    path = str(tmp_path / "big.holo")
    with holo.OutputHoloFile(path, holo.FrameDescriptor(1024, 1024, 2), 140) as out:
        out.meta_data = {"info": {"contiguous": 132}}
        out.write_header()
        out.write_frame(frames[:3])
        out.write_frame(frames[3:])
        out.write_footer()
        out.correct_number_of_frames(132)
    data = open(path, "rb").read()
    assert data == hf.holo_bytes(frames, 2, footer=holo.dump_footer({"info": {"contiguous": 132}}))


# ---- footers of header versions 1 to 6 --------------------------------------------------------

def _upgrades():
    with open(os.path.join(GOLDEN, "footer_upgrades.json")) as f:
        return json.load(f)


def test_footer_upgrades_match_reference_patches():
    """Every synthetic footer of header versions 1 to 8 upgraded as the
    reference's converter does with its own patch files (fixture written by
    tests/golden/make_footer_upgrades.py from those files): the same
    document, or a failure where the reference fails."""
    from holovibes_amd import footer_versions as fv
    cases = _upgrades()["footers"]
    assert len(cases) >= 19
    for c in cases:
        if c["expected"] is None:
            with pytest.raises(fv.FooterUpgradeError):
                fv.upgrade_footer(c["input"], c["version"])
        else:
            got = fv.upgrade_footer(c["input"], c["version"])
            assert got == c["expected"], c["case"]
    # the input is left as it was
    c = cases[0]
    before = json.dumps(c["input"], sort_keys=True)
    fv.upgrade_footer(c["input"], c["version"])
    assert json.dumps(c["input"], sort_keys=True) == before


def test_compute_settings_upgrade_matches_reference_patch():
    from holovibes_amd import footer_versions as fv
    for c in _upgrades()["settings"]:
        if c["expected"] is None:
            with pytest.raises(fv.FooterUpgradeError):
                fv.upgrade_compute_settings(c["input"])
        else:
            assert fv.upgrade_compute_settings(c["input"]) == c["expected"], c["case"]


def test_version3_recording_loads_its_settings(holo, tmp_path):
    """A recording with a version-3 header and footer: its settings come
    through the upgrade into hv_compute_settings (angular spectrum, STFT,
    T = 16, p = 3, z = 0.25 m, lambda 785 nm, fft shift, contrast range),
    fields the old footer lacks take the defaults the upgrade adds (batch
    size 1), and the pixel size comes from the upgraded info."""
    case = next(c for c in _upgrades()["footers"] if c["case"] == "v3 complete")
    frames = _frames(2, 16, 16, 2)
    path = str(tmp_path / "v3.holo")
    open(path, "wb").write(hf.holo_bytes(frames, 2, hf.RAW, footer=case["input"], version=3))
    with holo.InputHoloFile(path) as f:
        cs = f.import_compute_settings()
        assert cs == case["expected"]["compute_settings"]
        info = f.import_info()
    assert info == {"pixel_size": 20.0, "input_fps": 1, "camera_fps": 1}
    s, _ = holo.compute_settings_from_json(cs, holo.FrameDescriptor(16, 16, 2), info)
    from holovibes_amd.constants import HV_SPACE_ANGULARSP, HV_TIME_STFT
    assert (s.space_transformation, s.time_transformation, s.time_transformation_size) == \
        (HV_SPACE_ANGULARSP, HV_TIME_STFT, 16)
    assert (s.p_start, s.batch_size, s.time_stride, s.fft_shift) == (3, 1, 1, 1)
    assert s.z_distance == pytest.approx(0.25) and s.lambda_ == pytest.approx(7.85e-7)
    assert (s.contrast_min, s.contrast_max) == (0.25, 3.5) and s.pixel_size == 20.0
    assert s.renorm_enabled == 0


def test_unknown_enum_names_read_as_first_entry(holo):
    """nlohmann reads an enum name outside its table as the table's first
    entry: an unknown image type is MODULUS, an unknown time transformation
    STFT; a settings object without a version gets the v0 -> v1 edits (and
    fails without the keys they remove)."""
    from holovibes_amd import footer_versions as fv
    from holovibes_amd.constants import HV_IMG_MODULUS, HV_TIME_STFT
    cs = {"version": "v1", "image_rendering": {"image_mode": "HOLOGRAM", "time_transformation": "FFT9"},
          "view": {"image_type": "SQUAREDMODULUS"}}
    s, _ = holo.compute_settings_from_json(cs, holo.FrameDescriptor(16, 16, 2))
    assert (s.img_type, s.time_transformation) == (HV_IMG_MODULUS, HV_TIME_STFT)
    with pytest.raises(fv.FooterUpgradeError):
        holo.compute_settings_from_json({"image_rendering": {"image_mode": "HOLOGRAM"}}, holo.FrameDescriptor(16, 16, 2))
