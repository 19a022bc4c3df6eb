// The environment term of the sky: the luminance seen along a direction, from the sky-view
// LUT pair plus the lit ground where the direction hits it (the environment-luminance
// texel, AtmosphereRenderer.glsl:248-275), and cube-map addressing and filtering of
// RGBA16F cube maps with mip chains.
#pragma once

#include <hip/hip_fp16.h>

#include "atmosphere_device.h"

namespace skyhip {

// The atmosphere renderer's frame for sky-view lookups: the block's basis vectors and
// positions, the camera radius and altitude, and the sky-view parameterisation's horizon
// angles (GetHorizonDownAngleFromR, AtmosphereRenderer.glsl:74-78).
struct EnvFrame {
    F3 up, front, sun, cam, ec;
    float r, alt0, horizon_up, horizon_down;
};

// The horizon angles are computed in FP64 on the host: acos(sqrt(alt (r + R_b)) / r) is
// ill-conditioned near r = R_b.
inline EnvFrame make_env_frame(const SkyhipAtmosphereRenderBufferData& render, const SkyhipAtmosphereBufferData& atm) {
    EnvFrame e;
    e.up = f3(render.up_direction);
    e.front = f3(render.front_direction);
    e.sun = f3(render.sun_direction);
    e.cam = f3(render.camera_position);
    e.ec = f3(render.earth_center);
    const double r = render.camera_earth_center_distance, rb = atm.bottom_radius;
    const double alt0 = r - rb;
    const double down = std::acos(std::sqrt(std::fmax(alt0 * (r + rb), 0.0)) / r);
    e.r = float(r);
    e.alt0 = float(alt0);
    e.horizon_down = float(down);
    e.horizon_up = float(3.14159265358979323846 - down);
    return e;
}

// Luminance along world direction d (unit): GetCosLatLonFromViewDirection (:134-145),
// GetSkyViewTextureUvFromCosLatLon (:113-132) with bilinear CLAMP_TO_EDGE lookups of the
// sky-view luminance and transmittance, and ComputeGroundLuminance (Atmosphere.glsl:297-306)
// at the ground point when the direction hits the ground.
__device__ inline F3 env_luminance(const AtmDev& A, const EnvFrame& e, const float4* __restrict__ T, int tw, int th,
                                   const float4* __restrict__ sv_l, const float4* __restrict__ sv_t, int sw, int sh,
                                   F3 d) {
    const float cos_lat = dot(e.up, d);
    const F3 lon = d - e.up * cos_lat;
    const float l2 = dot(lon, lon);
    const float cos_lon = l2 == 0.0f ? 1.0f : dot(lon * rsqrtf(l2), e.front);
    const float lat = acosf(fminf(fmaxf(cos_lat, -1.0f), 1.0f));
    float x_lat;
    if (lat < e.horizon_up) {
        x_lat = 0.5f - 0.5f * sqrtf(fmaxf(1.0f - lat / e.horizon_up, 0.0f));
    } else {
        x_lat = sqrtf(fmaxf((lat - e.horizon_up) / e.horizon_down, 0.0f)) * 0.5f + 0.5f;
    }
    const float x_lon = sqrtf(fmaxf(0.5f - 0.5f * cos_lon, 0.0f));
    // Texel coordinates x (N - 1): GetTextureCoordFromUnitRange's uv = (0.5 + x (N - 1)) / N.
    F3 lum = lut_bilinear(sv_l, sw, sh, x_lon * float(sw - 1), x_lat * float(sh - 1));
    if (ray_hits_ground(A, e.r, e.alt0, cos_lat)) {
        const F3 tr = lut_bilinear(sv_t, sw, sh, x_lon * float(sw - 1), x_lat * float(sh - 1));
        const float dist = distance_to_bottom(A, e.r, e.alt0, cos_lat);
        F3 n = (e.cam - e.ec) + d * dist;
        n = n * (1.0f / sqrtf(dot(n, n)));
        const float mu_s = dot(e.sun, n);
        const F3 vis = sun_visibility(A, T, tw, th, A.rb, 0.0f, 0.0f, mu_s) * A.solar;
        lum += (1.0f / kPiF) * clamp01(mu_s) * A.albedo * vis * tr;
    }
    return lum;
}

// ConvertCubUvToDir (shaders/Base/Common.glsl:13-30): face index and face uv in [0, 1] to the
// (normalised) direction.
__host__ __device__ inline F3 cube_uv_to_dir(int face, float u, float v) {
    const float uc = 2.0f * u - 1.0f, vc = 2.0f * v - 1.0f;
    F3 d;
    switch (face) {
        case 0: d = {1.0f, vc, -uc}; break;
        case 1: d = {-1.0f, vc, uc}; break;
        case 2: d = {uc, 1.0f, -vc}; break;
        case 3: d = {uc, -1.0f, vc}; break;
        case 4: d = {uc, vc, 1.0f}; break;
        default: d = {-uc, vc, -1.0f}; break;
    }
    const float inv = 1.0f / sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    return d * inv;
}

// The direction of texel (x, y) of a w x w face as the environment and prefilter passes
// compute it: face_uv = ((x, y) + 0.5) / w with v flipped (AtmosphereRenderer.glsl:254-256).
__host__ __device__ inline F3 cube_texel_dir(int face, int x, int y, int w) {
    const float u = (float(x) + 0.5f) / float(w);
    const float v = 1.0f - (float(y) + 0.5f) / float(w);
    return cube_uv_to_dir(face, u, v);
}

// Cube-map face selection (the OpenGL cube-map face table): the major axis picks the face;
// ties go to x, then y.  s, t in [0, 1] address the face's texels (x = s w - 0.5 ...).
__device__ inline void cube_face_st(F3 d, int& face, float& s, float& t) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= 0.0f) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= 0.0f) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= 0.0f) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    }
    s = 0.5f * (sc / ma + 1.0f);
    t = 0.5f * (tc / ma + 1.0f);
}

// Texel offset of level l of a cube with base width w: levels are stored one after another,
// each as 6 faces of w_l x w_l RGBA16F texels, w_l = max(1, w >> l).
__host__ __device__ inline long long cube_level_offset(int w, int level) {
    long long off = 0;
    for (int l = 0; l < level; ++l) {
        const long long wl = w >> l > 0 ? w >> l : 1;
        off += 6 * wl * wl;
    }
    return off;
}

__device__ inline F3 rgba16f_rgb(const ushort4 v) {
    return {float(__builtin_bit_cast(_Float16, v.x)), float(__builtin_bit_cast(_Float16, v.y)),
            float(__builtin_bit_cast(_Float16, v.z))};
}

// Bilinear CLAMP_TO_EDGE lookup of one level at face coordinates (s, t); the filter footprint
// stays on the face (seamless filtering across face edges is not reproduced: it differs
// only within half a texel of an edge).
__device__ inline F3 cube_level_bilinear(const ushort4* __restrict__ cube, int w, int level, int face, float s,
                                         float t) {
    const int wl = max(w >> level, 1);
    const ushort4* f = cube + cube_level_offset(w, level) + (long long)face * wl * wl;
    const float fx = fminf(fmaxf(s * float(wl) - 0.5f, 0.0f), float(wl - 1));
    const float fy = fminf(fmaxf(t * float(wl) - 0.5f, 0.0f), float(wl - 1));
    const int x0 = int(fx), y0 = int(fy);
    const int x1 = min(x0 + 1, wl - 1), y1 = min(y0 + 1, wl - 1);
    const float ax = fx - float(x0), ay = fy - float(y0), bx = 1.0f - ax, by = 1.0f - ay;
    const F3 a = rgba16f_rgb(f[y0 * wl + x0]), b = rgba16f_rgb(f[y0 * wl + x1]);
    const F3 c = rgba16f_rgb(f[y1 * wl + x0]), dd = rgba16f_rgb(f[y1 * wl + x1]);
    return (a * bx + b * ax) * by + (c * bx + dd * ax) * ay;
}

// textureLod with LINEAR magnification and LINEAR_MIPMAP_LINEAR minification (the
// anisotropy sampler with an explicit level of detail is isotropic): lod <= 0 samples
// level 0; otherwise the two levels around lod, clamped to the last one, are blended.
__device__ inline F3 cube_sample_lod(const ushort4* __restrict__ cube, int w, int levels, F3 d, float lod) {
    int face;
    float s, t;
    cube_face_st(d, face, s, t);
    if (!(lod > 0.0f)) return cube_level_bilinear(cube, w, 0, face, s, t);
    const float lmax = float(levels - 1);
    lod = fminf(lod, lmax);
    const int l0 = int(lod);
    const int l1 = min(l0 + 1, levels - 1);
    const float a = lod - float(l0);
    const F3 c0 = cube_level_bilinear(cube, w, l0, face, s, t);
    if (a == 0.0f) return c0;
    const F3 c1 = cube_level_bilinear(cube, w, l1, face, s, t);
    return c0 * (1.0f - a) + c1 * a;
}

}  // namespace skyhip
