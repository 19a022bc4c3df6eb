// Test infrastructure (the oracle): CPU restatement of the cloud-texture noise
// generators, shaders/Base/Noise.glsl and shaders/SkyRendering/NoiseGen.comp,
// plus the UNORM8 store and the mip chain this build defines for
// glGenerateTextureMipmap (2x2 / 2x2x2 box, round half up).
#pragma once

#include "oracle_common.h"

namespace oracle {

// Noise.glsl:1-9
inline uint32_t WangHash(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

// Noise.glsl:11-15
inline uint32_t PCGHash(uint32_t seed) {
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Noise.glsl:17-22
constexpr int kPerlinGradients[16][3] = {{1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1},
                                         {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1},  {0, 1, -1}, {0, -1, -1},
                                         {1, 1, 0},  {-1, 1, 0},  {0, -1, 1}, {0, -1, -1}};

// Noise.glsl:24-26
inline const int* GetPerlinGradients(uint32_t i, uint32_t j, uint32_t k, uint32_t seed) {
    return kPerlinGradients[WangHash(seed + WangHash(i + WangHash(j + WangHash(k)))) & 0xfu];
}

// uvec3(ivec3(x)): float -> int -> uint (two's-complement wrap for negatives).
template <class S> inline uint32_t f2i2u(S x) { return uint32_t(int32_t(x)); }

// Noise.glsl:28-60
template <class S>
S PerlinNoise(V3<S> p, uint32_t freq, uint32_t seed) {
    p = p * S(freq);
    V3<S> fl = {std::floor(p.x), std::floor(p.y), std::floor(p.z)};
    V3<S> ce = {std::ceil(p.x), std::ceil(p.y), std::ceil(p.z)};
    uint32_t i0 = f2i2u(fl.x) % freq, j0 = f2i2u(fl.y) % freq, k0 = f2i2u(fl.z) % freq;
    uint32_t i1 = f2i2u(ce.x) % freq, j1 = f2i2u(ce.y) % freq, k1 = f2i2u(ce.z) % freq;
    V3<S> t = p - fl;
    auto fade = [](S x) { return x * x * x * (x * (x * S(6) - S(15)) + S(10)); };
    S u = fade(t.x), v = fade(t.y), w = fade(t.z);
    S x0 = t.x, y0 = t.y, z0 = t.z, x1 = t.x - S(1), y1 = t.y - S(1), z1 = t.z - S(1);
    auto g = [&](uint32_t i, uint32_t j, uint32_t k, S x, S y, S z) {
        const int* gr = GetPerlinGradients(i, j, k, seed);
        return S(gr[0]) * x + S(gr[1]) * y + S(gr[2]) * z;
    };
    return gmix(gmix(gmix(g(i0, j0, k0, x0, y0, z0), g(i1, j0, k0, x1, y0, z0), u),
                     gmix(g(i0, j1, k0, x0, y1, z0), g(i1, j1, k0, x1, y1, z0), u), v),
                gmix(gmix(g(i0, j0, k1, x0, y0, z1), g(i1, j0, k1, x1, y0, z1), u),
                     gmix(g(i0, j1, k1, x0, y1, z1), g(i1, j1, k1, x1, y1, z1), u), v),
                w);
}

// Noise.glsl:81-101 (3D Worley; NoiseGen.comp always passes a vec3)
template <class S>
S WorleyNoise(V3<S> p, uint32_t freq, uint32_t seed) {
    p = p * S(freq);
    uint32_t ijk[3] = {uint32_t(std::floor(p.x)), uint32_t(std::floor(p.y)), uint32_t(std::floor(p.z))};
    p = p + v3(S(freq));
    S min_dist = S(1e10);
    for (uint32_t di = freq - 1; di <= freq + 1; ++di)
        for (uint32_t dj = freq - 1; dj <= freq + 1; ++dj)
            for (uint32_t dk = freq - 1; dk <= freq + 1; ++dk) {
                uint32_t gx = ijk[0] + di, gy = ijk[1] + dj, gz = ijk[2] + dk;
                uint32_t sx = gx % freq, sy = gy % freq, sz = gz % freq;
                uint32_t rnd0 = WangHash(seed + WangHash(sx + WangHash(sy + WangHash(sz))));
                uint32_t rnd1 = WangHash(seed + WangHash(sx + WangHash(sy + WangHash(sz + 1))));
                uint32_t rnd2 = WangHash(seed + WangHash(sx + WangHash(sy + WangHash(sz + 2))));
                V3<S> gp = {S(gx) + S(rnd0) / S(4294967296.0), S(gy) + S(rnd1) / S(4294967296.0),
                            S(gz) + S(rnd2) / S(4294967296.0)};
                min_dist = std::fmin(min_dist, distance(p, gp));
            }
    return min_dist;
}

// NoiseGen.comp:12-18
template <class S> S RemapTo01(S x, S x0, S x1) { return gclamp((x - x0) / (x1 - x0), S(0), S(1)); }
template <class S> S RemapFrom01(S x, S y0, S y1) { return gclamp(y0 + x * (y1 - y0), S(0), S(1)); }

// NoiseGen.comp:20-34
template <class S>
S PerlinFBM(V3<S> p, const SkyhipNoiseCreateInfo& ci) {
    S res = 0;
    uint32_t f = ci.base_frequency;
    S a = S(0.5), sum_a = 0;
    for (int c = 0; c < 8; ++c) {
        S noise = PerlinNoise(p, f, ci.seed) * S(0.5) + S(0.5);
        res += RemapTo01(noise, S(ci.remap_min), S(ci.remap_max)) * a;
        sum_a += a;
        f *= 2;
        a *= S(0.5);
    }
    return res / sum_a;
}

// NoiseGen.comp:36-50
template <class S>
S WorleyFBM(V3<S> p, const SkyhipNoiseCreateInfo& ci) {
    S res = 0;
    uint32_t f = ci.base_frequency;
    S a = S(0.5), sum_a = 0;
    for (int c = 0; c < 8; ++c) {
        S noise = WorleyNoise(p, f, ci.seed);
        res += RemapTo01(noise, S(ci.remap_min), S(ci.remap_max)) * a;
        sum_a += a;
        f *= 2;
        a *= S(0.5);
    }
    return res / sum_a;
}

// UNORM8 conversion on imageStore: round(clamp(f, 0, 1) * 255), halves up.
template <class S> inline uint8_t to_unorm8(S f) {
    S c = gclamp(f, S(0), S(1)) * S(255);
    return uint8_t(std::floor(c + S(0.5)));
}

// NoiseGen.comp CLOUD_MAP_GEN (:52-67): texel -> (density, height) as UNORM8 RG.
template <class S>
void CloudMapTexel(int x, int y, int size, const SkyhipNoiseCreateInfo& density, const SkyhipNoiseCreateInfo& height,
                   uint8_t out[2]) {
    V3<S> coord = {(S(x) + S(0.5)) / S(size), (S(y) + S(0.5)) / S(size), S(0)};
    out[0] = to_unorm8(PerlinFBM(coord, density));
    out[1] = to_unorm8(WorleyFBM(coord, height));
}

// NoiseGen.comp DETAIL_MAP_GEN (:69-86): texel -> R8.
template <class S>
uint8_t DetailTexel(int x, int y, int z, int size, const SkyhipNoiseCreateInfo& perlin,
                    const SkyhipNoiseCreateInfo& worley) {
    V3<S> coord = {(S(x) + S(0.5)) / S(size), (S(y) + S(0.5)) / S(size), (S(z) + S(0.5)) / S(size)};
    S p = PerlinFBM(coord, perlin);
    S w = WorleyFBM(coord, worley);
    return to_unorm8(RemapFrom01(p, w, S(1)));
}

// The Worley detail volume (no reference counterpart): WorleyFBM (NoiseGen.comp:36-50) alone at
// the texel centre of a size^3 volume -> R8.
template <class S>
uint8_t WorleyTexel(int x, int y, int z, int size, const SkyhipNoiseCreateInfo& worley) {
    V3<S> coord = {(S(x) + S(0.5)) / S(size), (S(y) + S(0.5)) / S(size), (S(z) + S(0.5)) / S(size)};
    return to_unorm8(WorleyFBM(coord, worley));
}

// NoiseGen.comp DISPLACEMENT_GEN (:88-101): texel -> RGBA8, seed + channel.
template <class S>
void DisplacementTexel(int x, int y, int size, const SkyhipNoiseCreateInfo& perlin, uint8_t out[4]) {
    V3<S> coord = {(S(x) + S(0.5)) / S(size), (S(y) + S(0.5)) / S(size), S(0)};
    for (int i = 0; i < 4; ++i) {
        SkyhipNoiseCreateInfo ci = perlin;
        ci.seed += uint32_t(i);
        out[i] = to_unorm8(PerlinFBM(coord, ci));
    }
}

// Mip level `lvl` -> `lvl + 1` of a `channels`-channel UNORM8 texture with
// dims d (2 or 3 axes): box over the 2 (x 2 (x 2)) source texels, source
// indices clamped to the level's size, (sum + n/2) / n.
inline void downsample_u8(const uint8_t* src, const int sd[3], uint8_t* dst, const int dd[3], int dims, int channels) {
    int n = dims == 3 ? 8 : 4;
    for (int z = 0; z < (dims == 3 ? dd[2] : 1); ++z)
        for (int y = 0; y < dd[1]; ++y)
            for (int x = 0; x < dd[0]; ++x)
                for (int c = 0; c < channels; ++c) {
                    int sum = 0;
                    for (int k = 0; k < n; ++k) {
                        int sx = std::min(2 * x + (k & 1), sd[0] - 1), sy = std::min(2 * y + ((k >> 1) & 1), sd[1] - 1);
                        int sz = dims == 3 ? std::min(2 * z + ((k >> 2) & 1), sd[2] - 1) : 0;
                        sum += src[((size_t(sz) * sd[1] + sy) * sd[0] + sx) * channels + c];
                    }
                    dst[((size_t(z) * dd[1] + y) * dd[0] + x) * channels + c] = uint8_t((sum + n / 2) / n);
                }
}

}  // namespace oracle
