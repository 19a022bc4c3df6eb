// Test infrastructure (the oracle): CPU restatement of the environment-luminance cube
// (AtmosphereRenderer.glsl:248-275 with AtmosphereRenderer.cpp:234-242) and of the
// image-based-lighting precomputation (src/Base/src/IBL.cpp:25-45,
// shaders/Base/EnvRadianceSH.comp, shaders/Base/PrefilterRadiance.comp), with the rules
// the library defines where OpenGL leaves them to the implementation:
//  - cube maps are RGBA16F (binary16, round to nearest even), levels stored one after
//    another, faces +X -X +Y -Y +Z -Z, texel (x, y) row-major;
//  - glGenerateTextureMipmap is a 2x2 box, ((a + b) + (c + d)) * 0.25 in FP32;
//  - face selection follows the OpenGL cube-map table with ties to x, then y;
//  - lookups are bilinear CLAMP_TO_EDGE within a face (no filtering across face edges),
//    textureLod blends the two levels around the level of detail (lod <= 0: level 0).
// Only the tests and the benchmark's CPU baseline use this code.
#pragma once

#include <vector>

#include "oracle_pathtrace.h"

namespace oracle {

// ConvertCubUvToDir (shaders/Base/Common.glsl:13-30) at texel (x, y) of a w x w face, with the
// passes' v flip (face_uv.y = 1 - face_uv.y).
template <class S>
V3<S> CubeTexelDir(int face, int x, int y, int w) {
    S u = (S(x) + S(0.5)) / S(w), v = S(1) - (S(y) + S(0.5)) / S(w);
    S uc = S(2) * u - S(1), vc = S(2) * v - S(1);
    V3<S> d;
    switch (face) {
        case 0: d = {S(1), vc, -uc}; break;
        case 1: d = {S(-1), vc, uc}; break;
        case 2: d = {uc, S(1), -vc}; break;
        case 3: d = {uc, S(-1), vc}; break;
        case 4: d = {uc, vc, S(1)}; break;
        default: d = {-uc, vc, S(-1)}; break;
    }
    return normalize(d);
}

// The OpenGL cube-map face table: major axis, then (sc, tc, ma); s, t = (sc / |ma| + 1) / 2.
template <class S>
void CubeFaceST(V3<S> d, int& face, S& s, S& t) {
    S ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z), sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= 0) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= 0) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= 0) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    }
    s = S(0.5) * (sc / ma + S(1));
    t = S(0.5) * (tc / ma + S(1));
}

inline size_t CubeLevelOffset(int w, int level) {
    size_t off = 0;
    for (int l = 0; l < level; ++l) {
        size_t wl = size_t(std::max(w >> l, 1));
        off += 6 * wl * wl;
    }
    return off;
}

inline int Log2Int(int w) {
    int l = 0;
    while ((1 << (l + 1)) <= w) ++l;
    return l;
}

// Level `level` from level - 1, in place (cube: uint16 RGBA per texel, all levels).
inline void CubeMipLevel(uint16_t* cube, int w, int level) {
    int wl = std::max(w >> level, 1), wp = std::max(w >> (level - 1), 1);
    const uint16_t* src = cube + 4 * CubeLevelOffset(w, level - 1);
    uint16_t* dst = cube + 4 * CubeLevelOffset(w, level);
    for (int f = 0; f < 6; ++f)
        for (int y = 0; y < wl; ++y)
            for (int x = 0; x < wl; ++x) {
                int x0 = std::min(2 * x, wp - 1), x1 = std::min(2 * x + 1, wp - 1);
                int y0 = std::min(2 * y, wp - 1), y1 = std::min(2 * y + 1, wp - 1);
                const uint16_t* fp = src + 4 * size_t(f) * wp * wp;
                for (int c = 0; c < 4; ++c) {
                    float a = half_to_float(fp[4 * (y0 * wp + x0) + c]), b = half_to_float(fp[4 * (y0 * wp + x1) + c]);
                    float cc = half_to_float(fp[4 * (y1 * wp + x0) + c]), d = half_to_float(fp[4 * (y1 * wp + x1) + c]);
                    dst[4 * ((size_t(f) * wl + y) * wl + x) + c] = float_to_half_rne(((a + b) + (cc + d)) * 0.25f);
                }
            }
}

// Bilinear CLAMP_TO_EDGE lookup of one level within a face.
template <class S>
V3<S> CubeLevelBilinear(const uint16_t* cube, int w, int level, int face, S s, S t) {
    int wl = std::max(w >> level, 1);
    const uint16_t* f = cube + 4 * (CubeLevelOffset(w, level) + size_t(face) * wl * wl);
    S fx = gclamp(s * S(wl) - S(0.5), S(0), S(wl - 1)), fy = gclamp(t * S(wl) - S(0.5), S(0), S(wl - 1));
    int x0 = int(fx), y0 = int(fy), x1 = std::min(x0 + 1, wl - 1), y1 = std::min(y0 + 1, wl - 1);
    S ax = fx - S(x0), ay = fy - S(y0);
    auto px = [&](int x, int y) {
        const uint16_t* p = f + 4 * (y * wl + x);
        return V3<S>{S(half_to_float(p[0])), S(half_to_float(p[1])), S(half_to_float(p[2]))};
    };
    return (px(x0, y0) * (S(1) - ax) + px(x1, y0) * ax) * (S(1) - ay) + (px(x0, y1) * (S(1) - ax) + px(x1, y1) * ax) * ay;
}

// textureLod with LINEAR / LINEAR_MIPMAP_LINEAR filtering at an explicit level of detail.
template <class S>
V3<S> CubeSampleLod(const uint16_t* cube, int w, int levels, V3<S> d, S lod) {
    int face;
    S s, t;
    CubeFaceST(d, face, s, t);
    if (!(lod > S(0))) return CubeLevelBilinear(cube, w, 0, face, s, t);
    lod = std::fmin(lod, S(levels - 1));
    int l0 = int(lod), l1 = std::min(l0 + 1, levels - 1);
    S a = lod - S(l0);
    V3<S> c0 = CubeLevelBilinear(cube, w, l0, face, s, t);
    if (a == S(0)) return c0;
    return c0 * (S(1) - a) + CubeLevelBilinear(cube, w, l1, face, s, t) * a;
}

// The environment-luminance cube with its mip chain (out: uint16 RGBA, all levels).
template <class S>
void EnvironmentCube(const Atmosphere<S>& A, const RenderBlock<S>& rb, const float* T, int tw, int th,
                     const float* sv_l, const float* sv_t, int sw, int sh, int w, uint16_t* out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < 6 * w * w; ++i) {
        int face = i / (w * w), y = (i / w) % w, x = i % w;
        V3<S> L = EnvironmentLuminance(A, rb, T, tw, th, sv_l, sv_t, sw, sh, CubeTexelDir<S>(face, x, y, w));
        out[4 * size_t(i) + 0] = float_to_half_rne(float(L.x));
        out[4 * size_t(i) + 1] = float_to_half_rne(float(L.y));
        out[4 * size_t(i) + 2] = float_to_half_rne(float(L.z));
        out[4 * size_t(i) + 3] = 0;
    }
    for (int l = 1; l <= Log2Int(w); ++l) CubeMipLevel(out, w, l);
}

// EnvRadianceSH.comp: Llm[index] for the 9 coefficients (out: 9 x RGBA, a = 0), each a sum of
// 1024 stratified samples by the shader's halving tree.
template <class S>
void EnvRadianceSH(const uint16_t* cube, int w, S* out) {
    const int levels = Log2Int(w) + 1;
    const S Y00 = S(0.282095f), Y1n = S(0.488603f), Y2n = S(1.092548f), Y20 = S(0.315392f), Y22 = S(0.546274f);
    for (int index = 0; index < 9; ++index) {
        std::vector<V3<S>> L(1024);
        for (int li = 0; li < 1024; ++li) {
            S unit_theta = (S(0.5) + S(li >> 5)) / S(32), unit_phi = (S(0.5) + S(li & 31)) / S(32);
            S cos_theta = S(1) - S(2) * unit_theta;
            S sin_theta = std::sqrt(gclamp(S(1) - cos_theta * cos_theta, S(0), S(1)));
            S phi = S(2) * S(kPi) * unit_phi;
            V3<S> dir = {std::cos(phi) * sin_theta, cos_theta, std::sin(phi) * sin_theta};
            V3<S> rad = CubeSampleLod(cube, w, levels, dir, S(0));
            S c[9] = {Y00, Y1n * dir.y, Y1n * dir.z, Y1n * dir.x, Y2n * dir.x * dir.y, Y2n * dir.y * dir.z,
                      Y20 * (S(3) * dir.z * dir.z - S(1)), Y2n * dir.x * dir.z, Y22 * (dir.x * dir.x - dir.y * dir.y)};
            L[li] = rad * (c[index] * (S(4) * S(kPi) / S(1024)));
        }
        for (int half = 512; half >= 1; half >>= 1)
            for (int li = 0; li < half; ++li) L[li] = L[li] + L[li + half];
        out[4 * index + 0] = L[0].x, out[4 * index + 1] = L[0].y, out[4 * index + 2] = L[0].z, out[4 * index + 3] = 0;
    }
}

inline uint32_t ReverseBits32(uint32_t b) {  // shaders/Base/Noise.glsl:104-111
    b = (b << 16) | (b >> 16);
    b = ((b & 0x00ff00ffu) << 8) | ((b & 0xff00ff00u) >> 8);
    b = ((b & 0x0f0f0f0fu) << 4) | ((b & 0xf0f0f0f0u) >> 4);
    b = ((b & 0x33333333u) << 2) | ((b & 0xccccccccu) >> 2);
    b = ((b & 0x55555555u) << 1) | ((b & 0xaaaaaaaau) >> 1);
    return b;
}

// PrefilterEnvMap (PrefilterRadiance.comp) at direction R, roughness r, for an output level
// of width wl (the shader's invSaTexel uses the size of the image being written).
template <class S>
V3<S> PrefilterEnvMap(const uint16_t* env, int ew, int elevels, S roughness, V3<S> R, int wl) {
    S a = roughness * roughness, a2 = a * a;
    V3<S> N = R, V = R;
    // uint(mix(1, 64, pow(roughness, 0.3))) in the shader's FP32
    float p = std::pow(float(roughness), 0.3f);
    uint32_t n = uint32_t(1.0f * (1.0f - p) + 64.0f * p);
    S inv_sa_texel = (S(6) * S(wl) * S(wl)) / (S(4) * S(kPi));
    V3<S> t0, t1;
    PathTracer<S>::onb(N, t0, t1);
    V3<S> color = v3(S(0));
    S total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        S q = S(i) / S(n);
        S e1 = q - std::floor(q), e2 = S(ReverseBits32(i)) * S(2.3283064365386963e-10);  // Hammersley
        S phi = S(2) * S(kPi) * e1;                                                       // ImportanceSampleGGX
        S cos_t = std::sqrt((S(1) - e2) / (S(1) + (a2 - S(1)) * e2));
        S sin_t = std::sqrt(S(1) - cos_t * cos_t);
        V3<S> H = t0 * (sin_t * std::cos(phi)) + t1 * (sin_t * std::sin(phi)) + N * cos_t;
        V3<S> L = H * (S(2) * dot(V, H)) - V;
        S NoL = gclamp(dot(N, L), S(0), S(1));
        if (NoL > S(0)) {
            S NoH = gclamp(dot(N, H), S(0), S(1)), HoV = gclamp(dot(H, V), S(0), S(1));
            S dg = (NoH * a2 - NoH) * NoH + S(1);  // D_GGX (BRDF.glsl:36-40)
            S D = a2 / (S(kPi) * dg * dg);
            S pdf = std::fmax(D * NoH / (S(4) * HoV), S(0.0001));
            S sa_sample = S(1) / std::fmax(S(n) * pdf, S(0.00001));
            S mip = roughness == S(0) ? S(0) : S(0.5) * std::log2(sa_sample * inv_sa_texel) + S(2.5);
            color += CubeSampleLod(env, ew, elevels, L, mip) * NoL;
            total += NoL;
        }
    }
    return color / total;
}

// The prefiltered radiance cube: level l at roughness l / (count - 1), width pw >> l.
template <class S>
void PrefilteredRadiance(const uint16_t* env, int ew, int pw, int count, uint16_t* out) {
    const int elevels = Log2Int(ew) + 1;
    for (int l = 0; l < count; ++l) {
        int wl = std::max(pw >> l, 1);
        S roughness = count > 1 ? S(float(l) / float(count - 1)) : S(0);
        uint16_t* dst = out + 4 * CubeLevelOffset(pw, l);
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < 6 * wl * wl; ++i) {
            int face = i / (wl * wl), y = (i / wl) % wl, x = i % wl;
            V3<S> c = PrefilterEnvMap(env, ew, elevels, roughness, CubeTexelDir<S>(face, x, y, wl), wl);
            dst[4 * size_t(i) + 0] = float_to_half_rne(float(c.x));
            dst[4 * size_t(i) + 1] = float_to_half_rne(float(c.y));
            dst[4 * size_t(i) + 2] = float_to_half_rne(float(c.z));
            dst[4 * size_t(i) + 3] = 0;
        }
    }
}

}  // namespace oracle
