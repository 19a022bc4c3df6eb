"""State files: the reference's `bin/config*.json` (AppWindow::Init / SaveConfig,
src/SkyRendering/AppWindow.cpp:30-48, 128-137).

The reference serialises its parameter classes through a field-table reflection layer
(external/cppreflection/include/reflection/serialization.h) onto rapidjson:

- an object is a JSON object with one member per declared field, written in the field
  table's order (a std::map, so names sorted byte-wise); reading looks each declared field
  up by name (the first member of that name), keeps the current value of a field that is
  absent and reports it ("Field \"x\" not found"), and ignores members it does not declare;
- float / double: any JSON number (read as double, then rounded to float); int: a JSON
  integer that fits int32; bool; string; enums by their enumerator name (magic_enum);
  glm vectors and fixed arrays: arrays of exactly their length;
- a polymorphic material (std::unique_ptr<IVolumetricCloudMaterial>) is
  {"type": typeid(*p).name(), "data": {...}} or null; the type name is looked up in the
  subclass factory table (IVolumetricCloudMaterial.cpp:7-12);
- a type mismatch, a wrong array length, an unknown enumerator or material type is an error
  (R_ASSERT throws std::runtime_error; here StateFormatError);
- the text may hold comments and trailing commas (kParseCommentsFlag |
  kParseTrailingCommasFlag); it is written by rapidjson's PrettyWriter: 4-space indent, one
  array element per line, no newline after the root, doubles in rapidjson's dtoa format.

`load_state(text)` returns an AppState record tree with every field, `dump_state(state)` the
text SaveConfig would write for it, and the `to_*` helpers build this library's parameter
objects from a loaded state.
"""
from __future__ import annotations

import json
import math
import sys
from dataclasses import dataclass

import numpy as np


class StateFormatError(ValueError):
    """A state file that the reference's loader rejects (R_ASSERT / factory lookup)."""


# ---------------------------------------------------------------------------------------------
# Field types
# ---------------------------------------------------------------------------------------------

@dataclass(frozen=True)
class _Float:
    default: float

    def read(self, v, path):
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise StateFormatError(f"{path}: expected a number")
        return float(np.float32(v))

    def write(self, v):
        return float(np.float32(v))

    def fresh(self):
        return float(np.float32(self.default))


@dataclass(frozen=True)
class _Int:
    default: int

    def read(self, v, path):
        if isinstance(v, bool) or not isinstance(v, int) or not (-2 ** 31 <= v < 2 ** 31):
            raise StateFormatError(f"{path}: expected an int32 integer")
        return v

    def write(self, v):
        return int(v)

    def fresh(self):
        return self.default


@dataclass(frozen=True)
class _Bool:
    default: bool

    def read(self, v, path):
        if not isinstance(v, bool):
            raise StateFormatError(f"{path}: expected true or false")
        return v

    def write(self, v):
        return bool(v)

    def fresh(self):
        return self.default


@dataclass(frozen=True)
class _Enum:
    members: tuple
    default: str

    def read(self, v, path):
        if not isinstance(v, str):
            raise StateFormatError(f"{path}: expected an enumerator name")
        if v not in self.members:
            raise StateFormatError(f"{path}: unknown enumerator {v!r} (one of {', '.join(self.members)})")
        return v

    def write(self, v):
        return v

    def fresh(self):
        return self.default


@dataclass(frozen=True)
class _Vec:
    n: int
    default: tuple

    def read(self, v, path):
        if not isinstance(v, list) or len(v) != self.n:
            raise StateFormatError(f"{path}: expected an array of {self.n} numbers")
        return tuple(_Float(0.0).read(x, f"{path}[{i}]") for i, x in enumerate(v))

    def write(self, v):
        return [float(np.float32(x)) for x in v]

    def fresh(self):
        return tuple(float(np.float32(x)) for x in self.default)


@dataclass(frozen=True)
class _Obj:
    cls: str

    def read(self, v, path):
        out = Record(self.cls)
        out.load(v, path)
        return out

    def write(self, v):
        return v.to_json()

    def fresh(self):
        return Record(self.cls)


@dataclass(frozen=True)
class _Material:
    """std::unique_ptr<IVolumetricCloudMaterial>: {"type": name, "data": {...}} or null."""

    def read(self, v, path):
        if v is None:
            return None
        if not isinstance(v, dict):
            raise StateFormatError(f"{path}: expected an object or null")
        name = _first(v, "type")
        if name is _MISSING or not isinstance(name, str):
            raise StateFormatError(f"{path}: a material needs a string \"type\"")
        cls = MATERIAL_TYPES.get(name) or _ITANIUM_NAMES.get(name)
        if cls is None:
            raise StateFormatError(f"{path}: unknown material type {name!r}")
        data = _first(v, "data")
        if data is _MISSING:
            raise StateFormatError(f"{path}: a material needs \"data\"")
        rec = Record(cls)
        rec.load(data, f"{path}.data")
        return rec

    def write(self, v):
        if v is None:
            return None
        return {"type": _MSVC_NAMES[v.cls], "data": v.to_json()}

    def fresh(self):
        return None


_MISSING = object()


def _first(obj: dict, key):
    """rapidjson's FindMember: the first member of that name."""
    return obj.get(key, _MISSING)


def _noise(prefix, seed, freq, lo, hi):
    return {f"{prefix}.seed": _Int(seed), f"{prefix}.base_frequency": _Int(freq),
            f"{prefix}.remap_min": _Float(lo), f"{prefix}.remap_max": _Float(hi)}


_MIPMAP_MIN = ("NEAREST_MIPMAP_NEAREST", "LINEAR_MIPMAP_NEAREST", "NEAREST_MIPMAP_LINEAR", "LINEAR_MIPMAP_LINEAR")

# Field tables with the reference's defaults (the member initialisers; AppWindow's camera as its
# constructor leaves it, AppWindow.cpp:20-23).
SCHEMAS = {
    "AppWindow": {  # AppWindow.h:40-84
        "earth_": _Obj("Earth"),
        "volumetric_cloud_": _Obj("VolumetricCloud"),
        "camera_": _Obj("Camera"),
        "atmosphere_render_init_parameters_": _Obj("AtmosphereRenderInitParameters"),
        "atmosphere_render_parameters_": _Obj("AtmosphereRenderParameters"),
        "post_process_parameters_": _Obj("PostProcessParameters"),
        "camera_speed_": _Float(1.0),
        "draw_gui_enable_": _Bool(True),
        "draw_debug_textures_enable_": _Bool(False),
        "draw_help_enable_": _Bool(True),
        "full_screen_": _Bool(False),
        "anisotropy_enable_": _Bool(True),
        "vsync_enable_": _Bool(False),
        "smaa_option_": _Enum(("OFF", "SMAA_PRESET_LOW", "SMAA_PRESET_MEDIUM", "SMAA_PRESET_HIGH",
                               "SMAA_PRESET_ULTRA"), "SMAA_PRESET_HIGH"),  # SMAA.h:6-12
    },
    "Earth": {"parameters": _Obj("AtmosphereParameters"), "moon_status": _Obj("MoonStatus")},  # Earth.h:12-31
    "MoonStatus": {"direction_theta": _Float(70.0), "direction_phi": _Float(150.0),
                   "distance": _Float(384401.0 / 10.0), "radius": _Float(1737.0)},
    "AtmosphereParameters": {  # Atmosphere.h:10-67
        "solar_illuminance": _Vec(3, (1.0, 1.0, 1.0)),
        "sun_angular_radius": _Float(0.5334 * 0.5),
        "bottom_radius": _Float(6360.0),
        "thickness": _Float(6420.0 - 6360.0),
        "ground_albedo": _Vec(3, (0.4, 0.4, 0.4)),
        "rayleigh_exponential_distribution": _Float(8.0),
        "rayleigh_scattering_scale": _Float(0.0331),
        "rayleigh_scattering": _Vec(3, (0.175287, 0.409607, 1.0)),
        "mie_exponential_distribution": _Float(1.2),
        "mie_phase_g": _Float(0.8),
        "mie_scattering_scale": _Float(0.003996),
        "mie_scattering": _Vec(3, (1.0, 1.0, 1.0)),
        "mie_absorption_scale": _Float(0.000444),
        "mie_absorption": _Vec(3, (1.0, 1.0, 1.0)),
        "ozone_center_altitude": _Float(25.0),
        "ozone_width": _Float(15.0),
        "ozone_absorption_scale": _Float(0.001881),
        "ozone_absorption": _Vec(3, (0.345561, 1.0, 0.045189)),
        "transmittance_steps": _Float(40.0),
        "multiscattering_steps": _Float(30.0),
        "multiscattering_mask": _Float(1.0),
    },
    "Camera": {  # Camera.h:38-66
        "fovy": _Float(45.0), "zNear": _Float(3e-1), "zFar": _Float(5e4),
        "position_": _Vec(3, (0.0, 1.0, -1.0)),
        "front_": _Vec(3, (0.0, 0.0, -1.0)), "right_": _Vec(3, (1.0, 0.0, 0.0)), "up_": _Vec(3, (0.0, 1.0, 0.0)),
        "yaw_": _Float(180.0), "pitch_": _Float(0.0), "aspect_": _Float(1.0),
    },
    "AtmosphereRenderInitParameters": {  # AtmosphereRenderer.h:44-66
        "pcss_enable": _Bool(True), "volumetric_light_enable": _Bool(True), "moon_shadow_enable": _Bool(False),
        "raymarching_dither_sample_point_enable": _Bool(True), "use_sky_view_lut": _Bool(False),
        "sky_view_lut_dither_sample_point_enable": _Bool(False), "use_aerial_perspective_lut": _Bool(False),
        "aerial_perspective_lut_dither_sample_point_enable": _Bool(False),
        "aerial_perspective_lut_depth": _Int(32),
    },
    "AtmosphereRenderParameters": {  # AtmosphereRenderer.h:12-28
        "sun_direction_theta": _Float(70.0), "sun_direction_phi": _Float(180.0),
        "star_luminance_scale": _Float(0.005), "raymarching_steps": _Float(40.0),
        "sky_view_lut_steps": _Float(40.0), "aerial_perspective_lut_steps": _Float(40.0),
        "aerial_perspective_lut_max_distance": _Float(100.0),
    },
    "PostProcessParameters": {  # HDRBuffer.h:11-34
        "tone_mapping": _Enum(("CEToneMapping", "ACESToneMapping"), "ACESToneMapping"),
        "bloom_min_luminance": _Float(1.0), "bloom_max_delta_luminance": _Float(20.0),
        "bloom_filter_width": _Float(0.03), "bloom_intensity": _Float(0.1), "exposure": _Float(10.0),
        "dither_color_enable": _Bool(True),
    },
    "VolumetricCloud": {  # VolumetricCloud.h:16-56 (VolumetricCloud() leaves material null)
        "material": _Material(),
        "bottom_altitude_": _Float(2.0), "thickness_": _Float(2.0), "max_raymarch_distance_": _Float(30.0),
        "max_raymarch_steps_": _Float(128.0), "max_visible_distance_": _Float(120.0),
        "env_color_": _Vec(3, (1.0, 1.0, 1.0)), "env_color_scale_": _Float(0.1),
        "sun_illuminance_scale_": _Float(1.0), "shadow_steps_": _Float(5.0), "shadow_distance_": _Float(2.0),
        "shadow_map_max_distance": _Float(20.0), "shadow_froxel_max_distance": _Float(20.0),
        "sun_multiscattering_sigma_scale": _Float(0.3), "env_multiscattering_sigma_scale": _Float(0.5),
        "env_bottom_visibility": _Float(0.4), "env_sun_height_curve_exp": _Float(1.0),
    },
    "VolumetricCloudDefaultMaterialCommon": {  # VolumetricCloudDefaultMaterial.h:66-140, .cpp:30-57
        **_noise("cloud_map_.buffer.uDensity", 0, 3, 0.35, 0.75),
        **_noise("cloud_map_.buffer.uHeight", 0, 5, 0.8, 0.4),
        **_noise("detail_.buffer.uPerlin", 0, 7, 0.23, 1.0),
        **_noise("detail_.buffer.uWorley", 0, 11, 1.0, 0.0),
        **_noise("displacement_.buffer.uPerlin", 0, 6, 0.25, 0.75),
        "cloud_map_.texture.repeat_size": _Float(18.99), "detail_.texture.repeat_size": _Float(5.33),
        "displacement_.texture.repeat_size": _Float(3.51),
        "density_": _Float(15.0), "lod_bias_": _Float(2.75), "wind_speed_": _Float(0.05),
        "detail_wind_magnify_": _Float(1.0),
        "minfilter2d_": _Enum(_MIPMAP_MIN, "NEAREST_MIPMAP_NEAREST"),  # Samplers.h:20-25
        "minfilter3d_": _Enum(_MIPMAP_MIN, "NEAREST_MIPMAP_NEAREST"),
        "minfilter_displacement_": _Enum(_MIPMAP_MIN, "NEAREST_MIPMAP_NEAREST"),
    },
    "VolumetricCloudDefaultMaterial0": {  # VolumetricCloudDefaultMaterial.h:142-169
        "materail_common_": _Obj("VolumetricCloudDefaultMaterialCommon"),
        "detail_param_": _Vec(2, (0.4, 0.0)), "displacement_scale_": _Float(1.0),
    },
    "VolumetricCloudDefaultMaterial1": {  # VolumetricCloudDefaultMaterial.h:171-207
        "materail_common_": _Obj("VolumetricCloudDefaultMaterialCommon"),
        "base_density_threshold_": _Float(0.4), "base_height_hardness_": _Float(6.0),
        "base_edge_hardness_": _Float(6.0), "detail_base_": _Float(0.67), "detail_scale_": _Float(1.86),
        "height_cut_": _Float(0.9), "edge_cut_": _Float(0.8),
    },
    "VolumetricCloudMinimalMaterial": {"density_": _Float(0.5)},  # VolumetricCloudMinimalMaterial.h:8-27
    "VolumetricCloudVoxelMaterial": {  # VolumetricCloudVoxelMaterial.h:10-40
        "lod_bias_": _Float(2.75), "density_": _Float(20.0), "base_": _Vec(2, (0.0, 0.0)),
        "width_": _Vec(2, (2.0, 2.0)),
    },
}

# The subclass factory table's keys as the reference's MSVC build writes them (typeid names),
# and the Itanium-ABI names a GCC / Clang build of it would write (accepted when reading).
MATERIAL_TYPES = {f"class {c}": c for c in ("VolumetricCloudDefaultMaterial0", "VolumetricCloudDefaultMaterial1",
                                             "VolumetricCloudMinimalMaterial", "VolumetricCloudVoxelMaterial")}
_MSVC_NAMES = {c: k for k, c in MATERIAL_TYPES.items()}
_ITANIUM_NAMES = {f"{len(c)}{c}": c for c in MATERIAL_TYPES.values()}


class Record:
    """One reflected object: attribute per declared field, initialised to the defaults."""

    def __init__(self, cls: str):
        object.__setattr__(self, "cls", cls)
        object.__setattr__(self, "fields", SCHEMAS[cls])
        object.__setattr__(self, "values", {name: t.fresh() for name, t in SCHEMAS[cls].items()})
        object.__setattr__(self, "missing", [])

    def __getattr__(self, name):
        values = object.__getattribute__(self, "values")
        if name in values:
            return values[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name not in self.fields:
            raise AttributeError(f"{self.cls} has no serialised field {name!r}")
        self.values[name] = value

    def __getitem__(self, name):
        return self.values[name]

    def __setitem__(self, name, value):
        self.__setattr__(name, value)

    def __eq__(self, other):
        return isinstance(other, Record) and self.cls == other.cls and self.values == other.values

    def __repr__(self):
        return f"Record({self.cls}, {self.values})"

    def load(self, obj, path):
        """ISerialization::Deserialize (serialization.h:367-380) into this record."""
        if not isinstance(obj, dict):
            raise StateFormatError(f"{path}: expected an object")
        for name in sorted(self.fields):
            v = _first(obj, name)
            if v is _MISSING:
                self.missing.append(f"{path}.{name}")
                continue
            self.values[name] = self.fields[name].read(v, f"{path}.{name}")

    def to_json(self) -> dict:
        return {name: self.fields[name].write(self.values[name]) for name in sorted(self.fields)}

    def missing_fields(self) -> list:
        """Paths of the declared fields the loaded text did not hold, depth first."""
        out = list(self.missing)
        for name in sorted(self.fields):
            v = self.values[name]
            if isinstance(v, Record):
                out += v.missing_fields()
        return out


# ---------------------------------------------------------------------------------------------
# Text: rapidjson's reader flags and PrettyWriter
# ---------------------------------------------------------------------------------------------

def _strip_comments_and_trailing_commas(text: str) -> str:
    """Removes // and /* */ comments outside strings and commas that directly precede } or ]
    (with only whitespace or comments between)."""
    out = []
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            if j < 0:
                raise StateFormatError("unterminated /* comment")
            out.append(" ")
            i = j + 2
        else:
            out.append(c)
            i += 1
    s = "".join(out)
    # trailing commas: a comma followed by whitespace and then } or ] (outside strings)
    res = []
    i, n = 0, len(s)
    while i < n:
        c = s[i]
        if c == '"':
            j = i + 1
            while j < n and s[j] != '"':
                j += 2 if s[j] == "\\" else 1
            res.append(s[i:j + 1])
            i = j + 1
            continue
        if c == ",":
            j = i + 1
            while j < n and s[j] in " \t\r\n":
                j += 1
            if j < n and s[j] in "}]":
                i += 1
                continue
        res.append(c)
        i += 1
    return "".join(res)


def parse_json(text: str):
    """The document as rapidjson reads it with kParseCommentsFlag | kParseTrailingCommasFlag:
    objects keep their members' order and, for a repeated name, the first value wins."""
    def first_wins(pairs):
        d = {}
        for k, v in pairs:
            d.setdefault(k, v)
        return d

    def no_constants(name):
        raise StateFormatError(f"{name} is not JSON (the reference's reader rejects it)")

    try:
        return json.loads(_strip_comments_and_trailing_commas(text), object_pairs_hook=first_wins,
                          parse_constant=no_constants)
    except json.JSONDecodeError as e:
        raise StateFormatError(f"parse error at offset {e.pos}: {e.msg}") from None


_U64 = (1 << 64) - 1


def _cached_powers():
    """10^(-348 + 8 i), i = 0..86, as (f, e) with 2^63 <= f < 2^64 and f 2^e the power
    rounded to nearest: the cached-power table of Grisu (Loitsch 2010), computed exactly."""
    from fractions import Fraction
    out = []
    for i in range(87):
        v = Fraction(10) ** (-348 + 8 * i)
        e = v.numerator.bit_length() - v.denominator.bit_length() - 64
        while v / Fraction(2) ** e >= 2 ** 64:
            e += 1
        while v / Fraction(2) ** e < 2 ** 63:
            e -= 1
        f = math.floor(v / Fraction(2) ** e + Fraction(1, 2))
        if f == 2 ** 64:
            f, e = f >> 1, e + 1
        out.append((f, e))
    return out


_POWERS = None


def _mul(a, b):
    """The upper 64 bits of a 64 x 64-bit product, rounded half up on bit 63 of the lower half."""
    (fa, ea), (fb, eb) = a, b
    p = fa * fb
    h = p >> 64
    if (p >> 63) & 1:
        h += 1
    return h & _U64, ea + eb + 64


def _grisu2(v: float):
    """Grisu2 (Loitsch, "Printing floating-point numbers quickly and accurately with
    integers", PLDI 2010) as rapidjson's internal/dtoa.h runs it for v > 0: the digit string
    and its decimal exponent K (v ~ digits x 10^K).  The digits round-trip to v; they are the
    shortest in almost all cases, and where two shortest strings are equally near, Grisu2's
    choice can differ from a shortest-nearest-even printer (94.34500122070313)."""
    global _POWERS
    if _POWERS is None:
        _POWERS = _cached_powers()
    bits = int(np.float64(v).view(np.uint64))
    be, sig = (bits >> 52) & 0x7FF, bits & ((1 << 52) - 1)
    hidden = 1 << 52
    f, e = (sig + hidden, be - 1075) if be else (sig, -1074)
    # NormalizedBoundaries
    pf, pe = (f << 1) + 1, e - 1
    while not pf & (hidden << 1):
        pf, pe = pf << 1, pe - 1
    pf, pe = pf << 10, pe - 10
    mf, me = ((f << 2) - 1, e - 2) if f == hidden else ((f << 1) - 1, e - 1)
    mf, me = mf << (me - pe), pe
    # GetCachedPower
    dk = (-61 - pe) * 0.30102999566398114 + 347
    k = int(dk)
    if dk - k > 0.0:
        k += 1
    index = (k >> 3) + 1
    K = -(-348 + (index << 3))
    c = _POWERS[index]
    s = 64 - f.bit_length()
    W = _mul((f << s, e - s), c)
    Wp = _mul((pf, pe), c)
    Wm = _mul((mf, me), c)
    Wm = ((Wm[0] + 1) & _U64, Wm[1])
    Wp = ((Wp[0] - 1) & _U64, Wp[1])
    # DigitGen
    delta = (Wp[0] - Wm[0]) & _U64
    one_e = Wp[1]
    one_f = 1 << -one_e
    wp_w = (Wp[0] - W[0]) & _U64
    p1 = Wp[0] >> -one_e
    p2 = Wp[0] & (one_f - 1)
    buf = []

    def grisu_round(delta_, rest, ten_kappa, wpw):
        while (rest < wpw and ((delta_ - rest) & _U64) >= ten_kappa and
               ((rest + ten_kappa) & _U64 < wpw or ((wpw - rest) & _U64) > ((rest + ten_kappa - wpw) & _U64))):
            buf[-1] = chr(ord(buf[-1]) - 1)
            rest = (rest + ten_kappa) & _U64

    kappa = len(str(p1)) if p1 >= 10 else 1
    kappa = min(kappa, 9)
    while kappa > 0:
        div = 10 ** (kappa - 1)
        d, p1 = divmod(p1, div) if kappa > 1 else (p1, 0)
        if d or buf:
            buf.append(chr(ord("0") + d))
        kappa -= 1
        tmp = ((p1 << -one_e) + p2) & _U64
        if tmp <= delta:
            K += kappa
            grisu_round(delta, tmp, ((10 ** kappa) << -one_e) & _U64, wp_w)
            return "".join(buf), K
    while True:
        p2 = (p2 * 10) & _U64
        delta = (delta * 10) & _U64
        d = p2 >> -one_e
        if d or buf:
            buf.append(chr(ord("0") + d))
        p2 &= one_f - 1
        kappa -= 1
        if p2 < delta:
            K += kappa
            idx = -kappa
            grisu_round(delta, p2, one_f, (wp_w * (10 ** idx if idx < 20 else 0)) & _U64)
            return "".join(buf), K


def format_double(x: float) -> str:
    """rapidjson Writer::Double: Grisu2 digits, then internal/dtoa.h Prettify with
    maxDecimalPlaces 324: 1.0, 0.30000001192092896, 0.000444, 1e-7, 1.5e30."""
    if math.isnan(x) or math.isinf(x):
        raise StateFormatError("NaN and infinity are not written (the reference's writer refuses them)")
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    digits, k = _grisu2(abs(x))
    length = len(digits)
    kk = length + k  # 10^(kk-1) <= v < 10^kk
    if 0 <= k and kk <= 21:
        s = digits + "0" * k + ".0"
    elif 0 < kk <= 21:
        s = digits[:kk] + "." + digits[kk:]
    elif -6 < kk <= 0:
        s = "0." + "0" * (-kk) + digits
    elif kk < -324:
        s = "0.0"
    elif length == 1:
        s = digits + "e" + str(kk - 1)
    else:
        s = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + s


def dumps(value, indent: int = 0) -> str:
    """rapidjson PrettyWriter (4-space indent, kFormatDefault) of a JSON value."""
    pad = " " * (4 * (indent + 1))
    end = " " * (4 * indent)
    if value is None:
        return "null"
    if value is True:
        return "true"
    if value is False:
        return "false"
    if isinstance(value, int):
        return str(value)
    if isinstance(value, float):
        return format_double(value)
    if isinstance(value, str):
        return json.dumps(value, ensure_ascii=False)
    if isinstance(value, dict):
        if not value:
            return "{}"
        items = [f"{pad}{json.dumps(k, ensure_ascii=False)}: {dumps(v, indent + 1)}" for k, v in value.items()]
        return "{\n" + ",\n".join(items) + "\n" + end + "}"
    if isinstance(value, (list, tuple)):
        if not value:
            return "[]"
        return "[\n" + ",\n".join(pad + dumps(v, indent + 1) for v in value) + "\n" + end + "]"
    raise TypeError(f"cannot write {type(value).__name__}")


# ---------------------------------------------------------------------------------------------
# AppState: load / save
# ---------------------------------------------------------------------------------------------

def new_state() -> Record:
    """AppWindow's serialised state with every field at its default (material null)."""
    return Record("AppWindow")


def load_state(text: str, report_missing: bool = True) -> Record:
    """AppWindow::Init's Deserialize of a state file's text (AppWindow.cpp:30-44).  Absent
    fields keep their defaults and are listed by `missing_fields()` (and reported on stderr as
    the reference does when report_missing)."""
    state = new_state()
    state.load(parse_json(text), "")
    if report_missing:
        for p in state.missing_fields():
            print(f"Field \"{p.rsplit('.', 1)[-1]}\" not found", file=sys.stderr)
    return state


def load_state_file(path: str, report_missing: bool = True) -> Record:
    with open(path, encoding="utf-8") as f:
        return load_state(f.read(), report_missing)


def dump_state(state: Record) -> str:
    """AppWindow::SaveConfig's text (AppWindow.cpp:128-137) for `state`."""
    return dumps(state.to_json())


def save_state_file(state: Record, path: str):
    with open(path, "w", encoding="utf-8", newline="") as f:
        f.write(dump_state(state))


# ---------------------------------------------------------------------------------------------
# This library's parameter objects from a state
# ---------------------------------------------------------------------------------------------

def to_atmosphere_parameters(state: Record):
    from .atmosphere import AtmosphereParameters
    return AtmosphereParameters.from_json(state.earth_.parameters.to_json())


def to_camera(state: Record, aspect: float | None = None):
    """The state's camera.  The library's camera derives its basis from yaw and pitch
    (Camera::UpdateVectors); the stored front_/right_/up_ are that basis in the reference's
    files (a test checks it)."""
    from .atmosphere import Camera
    c = state.camera_
    return Camera(position=tuple(c.position_), yaw=c.yaw_, pitch=c.pitch_,
                  aspect=c.aspect_ if aspect is None else aspect, fovy=c.fovy, zNear=c.zNear, zFar=c.zFar)


def to_render_parameters(state: Record, camera=None):
    from .atmosphere import AtmosphereRenderParameters
    cam = camera or to_camera(state)
    r = state.atmosphere_render_parameters_
    return AtmosphereRenderParameters(
        sun_direction_theta=r.sun_direction_theta, sun_direction_phi=r.sun_direction_phi,
        star_luminance_scale=r.star_luminance_scale, raymarching_steps=r.raymarching_steps,
        sky_view_lut_steps=r.sky_view_lut_steps, aerial_perspective_lut_steps=r.aerial_perspective_lut_steps,
        aerial_perspective_lut_max_distance=r.aerial_perspective_lut_max_distance,
        camera_position=cam.position, view_projection=cam.view_projection())


def to_volumetric_cloud(state: Record):
    from .cloud import VolumetricCloud
    return VolumetricCloud.from_json(state.volumetric_cloud_.to_json())


def to_cloud_material(state: Record):
    """The host material object of the state's cloud material (None when it is null)."""
    from .cloud import (VolumetricCloudDefaultMaterial0, VolumetricCloudDefaultMaterial1,
                        VolumetricCloudMinimalMaterial, VolumetricCloudVoxelMaterial)
    m = state.volumetric_cloud_.material
    if m is None:
        return None
    data = m.to_json()
    if m.cls == "VolumetricCloudDefaultMaterial1":
        return VolumetricCloudDefaultMaterial1.from_json(data)
    if m.cls == "VolumetricCloudDefaultMaterial0":
        return VolumetricCloudDefaultMaterial0.from_json(data)
    if m.cls == "VolumetricCloudMinimalMaterial":
        return VolumetricCloudMinimalMaterial(density=m.density_)
    return VolumetricCloudVoxelMaterial(lod_bias=m.lod_bias_, density=m.density_, base=tuple(m.base_),
                                        width=tuple(m.width_))
