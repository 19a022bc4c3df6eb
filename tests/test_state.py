"""State files (bin/config*.json): load, save and the parameter objects built from them.

Golden vectors: the reference's own four state files (copied to tests/golden/configs), which
its SaveConfig wrote.  The saver must give them back byte for byte: key order (the field
tables' std::map), PrettyWriter layout and rapidjson's number format, including a Grisu2
digit choice that a shortest-nearest printer makes differently (94.34500122070313 in
config.json).  config_voxel.json lacks two cloud fields: they are reported, keep their
defaults, and are written back in their sorted places.  The loader's rules (first of
duplicate members, comments and trailing commas, type checks, unknown enumerators and
material types, the polymorphic {"type", "data"} form) are checked on edited copies.
"""
import json
import math
import os

import numpy as np
import pytest

from skyrendering_amd import _abi, presets, state
from skyrendering_amd.state import StateFormatError

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = ["config.json", "config2.json", "config3.json", "config_voxel.json"]


def _text(name):
    with open(os.path.join(HERE, "golden", "configs", name), encoding="utf-8") as f:
        return f.read()


@pytest.mark.parametrize("name", CONFIGS[:3])
def test_round_trip_is_byte_identical(name):
    text = _text(name)
    s = state.load_state(text, report_missing=False)
    assert s.missing_fields() == []
    assert state.dump_state(s) == text


def test_missing_fields_keep_defaults_and_are_written_back():
    text = _text("config_voxel.json")
    s = state.load_state(text, report_missing=False)
    assert s.missing_fields() == [".volumetric_cloud_.env_bottom_visibility",
                                  ".volumetric_cloud_.env_sun_height_curve_exp"]
    assert s.volumetric_cloud_.env_bottom_visibility == float(np.float32(0.4))
    assert s.volumetric_cloud_.env_sun_height_curve_exp == 1.0
    out = state.dump_state(s)
    lines, out_lines = text.split("\n"), out.split("\n")
    added = [l for l in out_lines if l not in lines]
    assert added == ['        "env_bottom_visibility": 0.4000000059604645,', '        "env_sun_height_curve_exp": 1.0,']
    assert [l for l in out_lines if l not in added] == lines
    assert state.dump_state(state.load_state(out, report_missing=False)) == out
    assert s.volumetric_cloud_.material.cls == "VolumetricCloudVoxelMaterial"


def test_reader_flags_duplicates_and_unknown_members():
    text = _text("config2.json")
    base = state.load_state(text, report_missing=False)
    # Comments and trailing commas (kParseCommentsFlag | kParseTrailingCommasFlag); a "//"
    # inside a string is not a comment.
    edited = text.replace('"camera_speed_":', '// speed\n    /* block\n comment */ "camera_speed_":', 1)
    edited = edited.replace('"vsync_enable_": false\n}', '"vsync_enable_": false,\n    "note": "a // b",\n}')
    s = state.load_state(edited, report_missing=False)
    assert s == base
    # The first of two members with one name wins; undeclared members are ignored.
    d = json.loads(text)
    dup = text.replace('"camera_speed_": ', '"camera_speed_": 77.0, "camera_speed_": ', 1)
    assert state.load_state(dup, report_missing=False).camera_speed_ == 77.0
    assert d["camera_speed_"] != 77.0


@pytest.mark.parametrize("old,new", [
    ('"camera_speed_": 29.0', '"camera_speed_": "fast"'),                # float from a string
    ('"aerial_perspective_lut_depth": 32', '"aerial_perspective_lut_depth": 32.0'),  # int from a double
    ('"aerial_perspective_lut_depth": 32', '"aerial_perspective_lut_depth": 4294967296'),  # beyond int32
    ('"anisotropy_enable_": true', '"anisotropy_enable_": 1'),          # bool from a number
    ('"smaa_option_": "SMAA_PRESET_HIGH"', '"smaa_option_": "SMAA_PRESET_EXTREME"'),  # unknown enumerator
    ('"type": "class VolumetricCloudDefaultMaterial1"', '"type": "class VolumetricCloudMaterial2"'),
    ('"camera_speed_": 29.0', '"camera_speed_": NaN'),
])
def test_loader_rejects_what_the_reference_asserts_on(old, new):
    text = _text("config.json")
    assert old in text
    with pytest.raises(StateFormatError):
        state.load_state(text.replace(old, new, 1), report_missing=False)
    # A vector of the wrong length.
    bad = text.replace('"env_color_": [\n            0.6372548937797546,', '"env_color_": [', 1)
    with pytest.raises(StateFormatError):
        state.load_state(bad, report_missing=False)


def test_polymorphic_material_forms():
    text = _text("config.json")
    s = state.load_state(text, report_missing=False)
    for cls in ("VolumetricCloudDefaultMaterial0", "VolumetricCloudDefaultMaterial1",
                "VolumetricCloudMinimalMaterial", "VolumetricCloudVoxelMaterial"):
        s.volumetric_cloud_.material = state.Record(cls)
        out = state.dump_state(s)
        assert f'"type": "class {cls}"' in out
        again = state.load_state(out, report_missing=False)
        assert again == s and state.dump_state(again) == out
    s.volumetric_cloud_.material = None
    out = state.dump_state(s)
    assert '"material": null' in out
    assert state.load_state(out, report_missing=False).volumetric_cloud_.material is None
    # A GCC / Clang build's typeid name is read and written back in the MSVC form.
    gcc = text.replace('"class VolumetricCloudDefaultMaterial1"', '"31VolumetricCloudDefaultMaterial1"')
    assert state.dump_state(state.load_state(gcc, report_missing=False)) == text


def test_double_format_kats_and_round_trip():
    kats = [(1.0, "1.0"), (100.0, "100.0"), (0.1, "0.1"), (0.000444, "0.000444"), (1e-7, "1e-7"),
            (1.5e30, "1.5e30"), (1e21, "1e21"), (123456789012345680000.0, "123456789012345680000.0"),
            (5e-324, "5e-324"), (float(np.float32(0.3)), "0.30000001192092896"), (-2.5, "-2.5"),
            (-0.0, "-0.0"), (0.0, "0.0"), (float(np.float32(94.345)), "94.34500122070313"),
            (1.7976931348623157e308, "1.7976931348623157e308")]
    for v, s in kats:
        assert state.format_double(v) == s, (v, state.format_double(v))
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 63 - 2 ** 52, 20000, dtype=np.int64).astype(np.uint64)
    vals = bits.view(np.float64)
    vals = np.concatenate([vals, rng.random(5000).astype(np.float32).astype(np.float64)])
    not_shortest = 0
    for v in vals:
        v = float(v)
        if math.isnan(v) or math.isinf(v):
            continue
        s = state.format_double(v)
        assert float(s) == v, (v, s)
        digits = lambda t: len(t.split("e")[0].replace("-", "").replace(".", "").strip("0"))
        not_shortest += digits(s) > digits(repr(v))
    assert not_shortest < 0.01 * len(vals)


def test_cached_powers_are_correctly_rounded():
    from fractions import Fraction
    powers = state._cached_powers()
    assert len(powers) == 87
    assert powers[0] == (0xFA8FD5A0081C0288, -1220)  # 10^-348, the published table's first entry
    for i, (f, e) in enumerate(powers):
        assert 2 ** 63 <= f < 2 ** 64
        err = abs(Fraction(f) * Fraction(2) ** e - Fraction(10) ** (-348 + 8 * i))
        assert err <= Fraction(2) ** e / 2


def test_parameter_objects_from_states():
    s = state.load_state(_text("config.json"), report_missing=False)
    atm = state.to_atmosphere_parameters(s)
    ref = presets.reference_earth()
    assert atm == ref or all(np.allclose(np.float32(getattr(atm, k)), np.float32(getattr(ref, k)))
                             for k in ref.__dataclass_fields__)
    cloud = state.to_volumetric_cloud(s)
    assert math.isclose(cloud.thickness, s.volumetric_cloud_.thickness_)
    m = state.to_cloud_material(s)
    assert m.kind == _abi.MATERIAL_DEFAULT1 and math.isclose(m.cloud_map_repeat_size, 94.34500122070312)
    v = state.to_cloud_material(state.load_state(_text("config_voxel.json"), report_missing=False))
    assert v.kind == _abi.MATERIAL_VOXEL and v.width == pytest.approx((6.918, 6.558), rel=1e-6)
    # The library's camera derives its basis from yaw and pitch (Camera::UpdateVectors): the
    # state files' stored front_ / right_ / up_ are that basis.
    for name in CONFIGS:
        st = state.load_state(_text(name), report_missing=False)
        cam = state.to_camera(st)
        front, right, up = cam.basis()
        assert np.allclose(front, st.camera_.front_, atol=2e-6), name
        assert np.allclose(right, st.camera_.right_, atol=2e-6), name
        assert np.allclose(up, st.camera_.up_, atol=2e-6), name
        rp = state.to_render_parameters(st, cam)
        assert rp.sun_direction_theta == st.atmosphere_render_parameters_.sun_direction_theta
