// Device-side cloud model shared by the raymarcher and the path tracer:
// UNORM8 texture sampling with the reference's sampler state, the material
// plugins' SampleSigmaT, the light march and the phase functions.
//
// Texture sampling is software, with exact GL semantics (GL 4.6, 8.14):
// textureLod with LINEAR magnification and NEAREST_MIPMAP_NEAREST
// minification: lambda <= 0.5 -> bilinear/trilinear on level 0, otherwise the
// nearest texel of level min(ceil(lambda + 0.5) - 1, q).  REPEAT wrapping for
// the default materials (power-of-two sizes), CLAMP_TO_BORDER(0) for voxels.
#pragma once

#include <hip/hip_fp16.h>

#include "atmosphere_device.h"

namespace skyhip {

constexpr int kMaxMipLevels = 16;

struct Tex2D {
    const uint8_t* data;
    int size, levels, channels;
    int log2_size;  // size is a power of two (REPEAT textures)
    uint32_t offset[kMaxMipLevels];
    const uint2* cells;  // RG8 level-0 cell-corner copy, or null
};

struct Tex3D {
    const uint8_t* data;
    int dim[3];
    int levels;
    int log2_size;  // cube edge log2 for REPEAT volumes (power-of-two cubes), -1 otherwise
    uint32_t offset[kMaxMipLevels];
    const uint2* cells;  // R8 level-0 cell-corner copy (REPEAT volumes), or null
};

// GL NEAREST_MIPMAP_NEAREST level for lambda > 0.5 (the caller handles
// lambda <= 0.5, including NaN and -inf, as magnification).
__device__ __forceinline__ int nearest_level(float lambda, int levels) {
    float d = ceilf(fminf(lambda, 64.0f) + 0.5f) - 1.0f;
    return min(int(d), levels - 1);
}

__device__ __forceinline__ int floor_to_int(float x) { return int(floorf(fminf(fmaxf(x, -1.0e9f), 1.0e9f))); }

// a + t (b - a): the linear interpolation GL's LINEAR filter weights form (exact in real arithmetic).
__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

// One texel of an RG8 (kChannels 2) or RGBA8 (4) texture as one 2- or 4-byte load; texel offsets are
// 32-bit and the base is uniform, so the load takes the scalar-base form.
template <int kChannels>
__device__ __forceinline__ uint32_t texel_bits(const uint8_t* base, uint32_t index) {
    const uint32_t byte_offset = index * uint32_t(kChannels);  // 32-bit offset from the uniform base
    if (kChannels == 2) return *reinterpret_cast<const uint16_t*>(base + byte_offset);
    return *reinterpret_cast<const uint32_t*>(base + byte_offset);
}
__device__ __forceinline__ float unorm_byte(uint32_t bits, int c) { return float((bits >> (8 * c)) & 0xffu); }

// textureLod(sampler2D REPEAT) of an RG8 / RGBA8 texture whose size is a power of two; out[c]
// for the first kChannels channels (t.channels == kChannels).
template <int kChannels>
__device__ inline void tex2d_repeat(const Tex2D& t, float u, float v, float lambda, float out[kChannels]) {
    if (!(lambda > 0.5f)) {
        const int s = t.size, ls = t.log2_size;
        const uint32_t m = uint32_t(s - 1);
        const float x = u * float(s) - 0.5f, y = v * float(s) - 0.5f;
        const int ix = floor_to_int(x), iy = floor_to_int(y);
        const float ax = x - floorf(x), ay = y - floorf(y);
        const uint32_t x0 = uint32_t(ix) & m, x1 = uint32_t(ix + 1) & m;
        const uint32_t r0 = (uint32_t(iy) & m) << ls, r1 = (uint32_t(iy + 1) & m) << ls;
        uint32_t b00, b10, b01, b11;
        if (kChannels == 2 && t.cells) {  // the cell's four RG8 texels in one 8-byte load
            const uint2 q = t.cells[r0 | x0];
            b00 = q.x & 0xffffu, b10 = q.x >> 16, b01 = q.y & 0xffffu, b11 = q.y >> 16;
        } else {
            b00 = texel_bits<kChannels>(t.data, r0 | x0), b10 = texel_bits<kChannels>(t.data, r0 | x1);
            b01 = texel_bits<kChannels>(t.data, r1 | x0), b11 = texel_bits<kChannels>(t.data, r1 | x1);
        }
        for (int c = 0; c < kChannels; ++c)
            out[c] = lerpf(lerpf(unorm_byte(b00, c), unorm_byte(b10, c), ax),
                           lerpf(unorm_byte(b01, c), unorm_byte(b11, c), ax), ay) * (1.0f / 255.0f);
        return;
    }
    const int l = nearest_level(lambda, t.levels);
    const int ls = max(t.log2_size - l, 0);
    const int s = 1 << ls;
    const uint32_t m = uint32_t(s - 1);
    const uint32_t xi = uint32_t(floor_to_int(u * float(s))) & m, yi = uint32_t(floor_to_int(v * float(s))) & m;
    const uint32_t b = texel_bits<kChannels>(t.data + t.offset[l], (yi << ls) | xi);
    for (int c = 0; c < kChannels; ++c) out[c] = unorm_byte(b, c) * (1.0f / 255.0f);
}

// textureLod(sampler3D REPEAT) of an R8 volume that is a power-of-two cube.
__device__ inline float tex3d_r8_repeat(const Tex3D& t, float u, float v, float w, float lambda) {
    const uint8_t* d = t.data;
    if (!(lambda > 0.5f)) {
        const int ls = t.log2_size, s = 1 << ls;
        const uint32_t m = uint32_t(s - 1);
        const float x = u * float(s) - 0.5f, y = v * float(s) - 0.5f, z = w * float(s) - 0.5f;
        const int ix = floor_to_int(x), iy = floor_to_int(y), iz = floor_to_int(z);
        const float ax = x - floorf(x), ay = y - floorf(y), az = z - floorf(z);
        const uint32_t x0 = uint32_t(ix) & m, y0 = (uint32_t(iy) & m) << ls, z0 = (uint32_t(iz) & m) << (2 * ls);
        float c00, c10, c01, c11;
        if (t.cells) {  // the cell's eight texels, byte a + 2b + 4c, in one 8-byte load
            const uint2 q = t.cells[z0 | y0 | x0];
            c00 = lerpf(float(q.x & 0xffu), float((q.x >> 8) & 0xffu), ax);
            c10 = lerpf(float((q.x >> 16) & 0xffu), float(q.x >> 24), ax);
            c01 = lerpf(float(q.y & 0xffu), float((q.y >> 8) & 0xffu), ax);
            c11 = lerpf(float((q.y >> 16) & 0xffu), float(q.y >> 24), ax);
        } else {
            const uint32_t x1 = uint32_t(ix + 1) & m, y1 = (uint32_t(iy + 1) & m) << ls,
                           z1 = (uint32_t(iz + 1) & m) << (2 * ls);
            c00 = lerpf(float(d[z0 | y0 | x0]), float(d[z0 | y0 | x1]), ax);
            c10 = lerpf(float(d[z0 | y1 | x0]), float(d[z0 | y1 | x1]), ax);
            c01 = lerpf(float(d[z1 | y0 | x0]), float(d[z1 | y0 | x1]), ax);
            c11 = lerpf(float(d[z1 | y1 | x0]), float(d[z1 | y1 | x1]), ax);
        }
        return lerpf(lerpf(c00, c10, ay), lerpf(c01, c11, ay), az) * (1.0f / 255.0f);
    }
    const int l = nearest_level(lambda, t.levels);
    const int ls = max(t.log2_size - l, 0);
    const uint32_t m = uint32_t((1 << ls) - 1);
    const float s = float(1 << ls);
    const uint32_t xi = uint32_t(floor_to_int(u * s)) & m, yi = uint32_t(floor_to_int(v * s)) & m,
                   zi = uint32_t(floor_to_int(w * s)) & m;
    return float(d[t.offset[l] + ((zi << (2 * ls)) | (yi << ls) | xi)]) * (1.0f / 255.0f);
}

// textureLod(sampler3D CLAMP_TO_BORDER(0)) of an R8 volume of any dimensions.
__device__ inline float tex3d_r8_border(const Tex3D& t, float u, float v, float w, float lambda) {
    const uint8_t* d = t.data;
    if (!(lambda > 0.5f)) {
        const int sx = t.dim[0], sy = t.dim[1], sz = t.dim[2];
        float x = u * float(sx) - 0.5f, y = v * float(sy) - 0.5f, z = w * float(sz) - 0.5f;
        int ix = floor_to_int(x), iy = floor_to_int(y), iz = floor_to_int(z);
        float ax = x - floorf(x), ay = y - floorf(y), az = z - floorf(z);
        float r = 0.0f;
        for (int k = 0; k < 8; ++k) {
            int xx = ix + (k & 1), yy = iy + ((k >> 1) & 1), zz = iz + ((k >> 2) & 1);
            float wgt = ((k & 1) ? ax : 1.0f - ax) * ((k & 2) ? ay : 1.0f - ay) * ((k & 4) ? az : 1.0f - az);
            bool inside = xx >= 0 && xx < sx && yy >= 0 && yy < sy && zz >= 0 && zz < sz;
            float texel = inside ? float(d[((long long)zz * sy + yy) * sx + xx]) : 0.0f;
            r += wgt * texel;
        }
        return r * (1.0f / 255.0f);
    }
    const int l = nearest_level(lambda, t.levels);
    const int sx = max(1, t.dim[0] >> l), sy = max(1, t.dim[1] >> l), sz = max(1, t.dim[2] >> l);
    int xi = floor_to_int(u * float(sx)), yi = floor_to_int(v * float(sy)), zi = floor_to_int(w * float(sz));
    if (xi < 0 || xi >= sx || yi < 0 || yi >= sy || zi < 0 || zi >= sz) return 0.0f;
    return float(d[t.offset[l] + ((long long)zi * sy + yi) * sx + xi]) * (1.0f / 255.0f);
}

// FP16 density grid (x fastest), sampled at level 0 with LINEAR filtering and
// CLAMP_TO_BORDER(0).
struct GridF16 {
    const uint16_t* data;
    int dim[3];
};

__device__ __forceinline__ float half_bits_to_float(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }

// Trilinear interpolation of a cell's corners c[a + 2b + 4c] at fractions (ax, ay, az): x, then
// y, then z, each as a lerp.
__device__ __forceinline__ float trilinear_corners(const float c[8], float ax, float ay, float az) {
    const float x0 = lerpf(c[0], c[1], ax), x1 = lerpf(c[2], c[3], ax);
    const float x2 = lerpf(c[4], c[5], ax), x3 = lerpf(c[6], c[7], ax);
    return lerpf(lerpf(x0, x1, ay), lerpf(x2, x3, ay), az);
}

// Trilinear sample at texel-space (x, y, z) = (u, v, w) * dim - 0.5 (GL's texel coordinates).
__device__ inline float grid_f16_trilinear_texel(const GridF16& g, float x, float y, float z) {
    const int nx = g.dim[0], ny = g.dim[1], nz = g.dim[2];
    int ix = floor_to_int(x), iy = floor_to_int(y), iz = floor_to_int(z);
    float ax = x - floorf(x), ay = y - floorf(y), az = z - floorf(z);
    float c[8];
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix + 1 < nx && iy + 1 < ny && iz + 1 < nz) {
        const uint16_t* p = g.data + ((long long)iz * ny + iy) * nx + ix;
        const long long plane = (long long)nx * ny;
        c[0] = half_bits_to_float(p[0]);
        c[1] = half_bits_to_float(p[1]);
        c[2] = half_bits_to_float(p[nx]);
        c[3] = half_bits_to_float(p[nx + 1]);
        c[4] = half_bits_to_float(p[plane]);
        c[5] = half_bits_to_float(p[plane + 1]);
        c[6] = half_bits_to_float(p[plane + nx]);
        c[7] = half_bits_to_float(p[plane + nx + 1]);
    } else {
        for (int k = 0; k < 8; ++k) {
            int xx = ix + (k & 1), yy = iy + ((k >> 1) & 1), zz = iz + ((k >> 2) & 1);
            bool inside = xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
            c[k] = inside ? half_bits_to_float(g.data[((long long)zz * ny + yy) * nx + xx]) : 0.0f;
        }
    }
    return trilinear_corners(c, ax, ay, az);
}

// The same sample from the grid's cell-corner copy (skyhip_voxel_grid_cells): cell (i, j, k),
// i in [-1, nx - 1], holds the eight texels (i + a, j + b, k + c) as FP16 in the order
// a + 2b + 4c, zeros outside the grid, cells x fastest from (-1, -1, -1).  One 16-byte load;
// samples whose cell lies outside the copy are 0, as all their texels are border texels.
__device__ __forceinline__ float grid_cells_trilinear_texel(const uint4* __restrict__ cells, const int dim[3], float x,
                                                            float y, float z) {
    const int ix = floor_to_int(x), iy = floor_to_int(y), iz = floor_to_int(z);
    const float ax = x - floorf(x), ay = y - floorf(y), az = z - floorf(z);
    const bool inside = ix >= -1 && iy >= -1 && iz >= -1 && ix < dim[0] && iy < dim[1] && iz < dim[2];
    const uint32_t cell = inside ? (uint32_t(iz + 1) * uint32_t(dim[1] + 1) + uint32_t(iy + 1)) *
                                           uint32_t(dim[0] + 1) + uint32_t(ix + 1)
                                 : 0u;
    const uint4 q = cells[cell];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
    float c[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[2 * k] = half_bits_to_float(uint16_t(w[k] & 0xffffu));
        c[2 * k + 1] = half_bits_to_float(uint16_t(w[k] >> 16));
    }
    const float r = trilinear_corners(c, ax, ay, az);
    return inside ? r : 0.0f;
}

// Trilinear sample at normalised (u, v, w).
__device__ inline float grid_f16_trilinear(const GridF16& g, float u, float v, float w) {
    return grid_f16_trilinear_texel(g, u * float(g.dim[0]) - 0.5f, v * float(g.dim[1]) - 0.5f,
                                    w * float(g.dim[2]) - 0.5f);
}

// Material parameters as the kernels see them.  lod_* = log2(k_lod) + uLodBias of each
// texture's SampleInfo, so that lambda = log2(k_lod distance) + bias is log2(distance) + lod_*
// with one log2 per density evaluation.
struct MaterialDev {
    int kind;
    float lod_cloud_map, lod_detail, lod_displacement, lod_worley;
    SkyhipCloudMaterialCommonBufferData common;
    SkyhipCloudMaterial1BufferData m1;
    SkyhipCloudMaterial0BufferData m0;
    float minimal_density;
    SkyhipCloudVoxelBufferData voxel;
    SkyhipCloudWorleyDetailBufferData worley;
    Tex2D cloud_map, displacement;
    Tex3D detail, voxel_tex, worley_tex;
    GridF16 grid;
};

__device__ __forceinline__ float length3(F3 a) { return sqrtf(dot(a, a)); }
// The hardware square root (v_sqrt_f32, within 1 ulp) where the result feeds a LOD or a
// height fraction: the correctly rounded sqrtf expands to about 20 instructions.
__device__ __forceinline__ float sqrt_1ulp(float x) { return __builtin_amdgcn_sqrtf(x); }

// Work tally of the counting kernel variants (per lane).
struct WorkCount {
    uint32_t evals = 0, detail = 0, cloud_steps = 0;
};

// SampleSigmaT(pos, height01) of the selected material (cam = uCameraPos,
// whose distance drives every LOD).
template <int kMat, bool kCount = false>
__device__ inline float sample_sigma_t(const MaterialDev& M, F3 pos, float height01, F3 cam,
                                       WorkCount* wc = nullptr) {
    if (kCount) wc->evals++;
    if (kMat == SKYHIP_MATERIAL_MINIMAL) return M.minimal_density;  // VolumetricCloudMaterialMinimal.glsl:6-7
    if (kMat == SKYHIP_MATERIAL_VOXEL_F16)  // the voxel material's uv mapping on the FP16 grid, level 0
        return grid_f16_trilinear(M.grid, pos.x * M.voxel.uSampleFrequency[0] + M.voxel.uSampleBias[0],
                                  pos.y * M.voxel.uSampleFrequency[1] + M.voxel.uSampleBias[1], height01) *
               M.voxel.uDensity;
    const F3 rel = pos - cam;
    const float dist = sqrt_1ulp(dot(rel, rel));  // LOD distance
    if (kMat == SKYHIP_MATERIAL_VOXEL) {  // VolumetricCloudMaterialVoxel.glsl:12-16
        float u = pos.x * M.voxel.uSampleFrequency[0] + M.voxel.uSampleBias[0];
        float v = pos.y * M.voxel.uSampleFrequency[1] + M.voxel.uSampleBias[1];
        float lod = __log2f(M.voxel.uSampleLodK * dist) + M.voxel.uLodBias;
        return tex3d_r8_border(M.voxel_tex, u, v, height01, lod) * M.voxel.uDensity;
    }
    // GetUVWLod (VolumetricCloudDefaultMaterialCommon.glsl:20-24) for every texture of the evaluation.
    const float log2_dist = __log2f(dist);
    const SkyhipSampleInfo& ci = M.common.uCloudMapSampleInfo;
    const SkyhipSampleInfo& di = M.common.uDetailSampleInfo;
    float ct[2];
    tex2d_repeat<2>(M.cloud_map, pos.x * ci.frequency + ci.bias[0], pos.y * ci.frequency + ci.bias[1],
                    log2_dist + M.lod_cloud_map, ct);
    if (kMat == SKYHIP_MATERIAL_DEFAULT1 || kMat == SKYHIP_MATERIAL_DEFAULT1_WORLEY) {
        // VolumetricCloudDefaultMaterial1.glsl:14-28
        const SkyhipCloudMaterial1BufferData& p = M.m1;
        float density = clamp01((ct[0] - p.uBaseDensityThreshold) * p.uBaseEdgeHardness);
        density *= clamp01((1.0f - height01) * p.uBaseHeightHardness);
        if (density == 0.0f) return 0.0f;
        if (kCount) wc->detail++;
        float detail = tex3d_r8_repeat(M.detail, pos.x * di.frequency + di.bias[0], pos.y * di.frequency + di.bias[1],
                                       pos.z * di.frequency, log2_dist + M.lod_detail);
        const float mask = fmaxf(clamp01(height01 - p.uHeightCut), clamp01(p.uEdgeCur - ct[0]));
        detail = (detail + p.uDetailBase) * p.uDetailScale;
        detail *= mask;
        density = clamp01(density - detail);
        if (kMat == SKYHIP_MATERIAL_DEFAULT1_WORLEY) {
            // The second erosion by the Worley volume (SkyhipCloudMaterialKind).  With a
            // non-negative base and scale the term is >= 0, so a zero density stays zero
            // without the fetch; otherwise the fetch is always made.
            const SkyhipCloudWorleyDetailBufferData& q = M.worley;
            const SkyhipSampleInfo& wi = q.uWorleySampleInfo;
            if (density > 0.0f || !(q.uWorleyBase >= 0.0f && q.uWorleyScale >= 0.0f)) {
                float w = tex3d_r8_repeat(M.worley_tex, pos.x * wi.frequency + wi.bias[0],
                                          pos.y * wi.frequency + wi.bias[1], pos.z * wi.frequency,
                                          log2_dist + M.lod_worley);
                w = (w + q.uWorleyBase) * q.uWorleyScale;
                density = clamp01(density - w * mask);
            }
        }
        return density * M.common.uDensity * height01;
    }
    // SKYHIP_MATERIAL_DEFAULT0, VolumetricCloudDefaultMaterial0.glsl:18-31
    const SkyhipSampleInfo& si = M.common.uDisplacementSampleInfo;
    float lod = log2_dist + M.lod_displacement;
    float u = pos.x * si.frequency + si.bias[0], v = pos.y * si.frequency + si.bias[1], w = pos.z * si.frequency;
    float d0[4], d1[4];
    tex2d_repeat<4>(M.displacement, u, v, lod, d0);
    tex2d_repeat<4>(M.displacement, u, w, lod, d1);
    F3 p2 = pos + M.m0.uDisplacementScale * F3{d0[0] + d1[2], d0[1], d1[3]};
    float detail = tex3d_r8_repeat(M.detail, p2.x * di.frequency + di.bias[0], p2.y * di.frequency + di.bias[1],
                                   p2.z * di.frequency, __log2f(length3(p2 - cam)) + M.lod_detail);
    detail = detail * M.m0.uDetailParam[0] + M.m0.uDetailParam[1];
    float hit = clamp01(height01 / ct[1]);                  // CalHeightMask (:9-12)
    float mask = clamp01(hit * (hit - 1.0f) * -4.0f);
    float x = ct[0] * mask;
    return clamp01((x - detail) / (1.0f - detail)) * height01 * M.common.uDensity;  // Remap01 (:14-16)
}

// HenyeyGreenstein, VolumetricCloudCommon.glsl:58-63.
__device__ __forceinline__ float henyey_greenstein(float c, float g) {
    float a = 1.0f - g * g;
    float b = 1.0f + g * g - 2.0f * g * c;
    return (0.25f / kPiF) * a / (b * sqrtf(b));
}

// Cloud-layer geometry in the local z-up frame (planet centre (0, 0, -R)).
struct Shell {
    F3 cam, sun;
    float R, hb, ht, inv_thickness;
};

// CalHeight01 (VolumetricCloudCommon.glsl:36-40) via |p'|^2 - R^2 = x^2 + y^2 + z (z + 2R).
__device__ __forceinline__ float height01(const Shell& S, F3 p, float* q_out = nullptr) {
    float q = p.x * p.x + p.y * p.y + p.z * (p.z + 2.0f * S.R);
    float alt = q * __builtin_amdgcn_rcpf(sqrt_1ulp(fmaxf(q + S.R * S.R, 0.0f)) + S.R);  // 1-ulp sqrt, reciprocal
    if (q_out) *q_out = q;
    return clamp01((alt - S.hb) * S.inv_thickness);
}

// SampleShadow, VolumetricCloudRender.comp:98-114: float counter, t^2 spacing.
template <int kMat, bool kCount = false>
__device__ inline float sample_shadow(const MaterialDev& M, const Shell& S, F3 pos, float shadow_steps,
                                      float shadow_distance, WorkCount* wc = nullptr) {
    float optical_depth = 0.0f;
    const float inv_steps = 1.0f / shadow_steps;
    const F3 sample_vector = S.sun * shadow_distance;
    float previous_t = 0.0f;
    for (float t = inv_steps; t <= 1.0f; t += inv_steps) {
        float current_t = t * t;
        float delta_t = current_t - previous_t;
        F3 sp = pos + sample_vector * (previous_t + 0.5f * delta_t);
        optical_depth += sample_sigma_t<kMat, kCount>(M, sp, height01(S, sp), S.cam, wc) * shadow_distance * delta_t;
        previous_t = current_t;
    }
    return __expf(-optical_depth);
}

// Roots of |o + t d|^2 = R'^2 for a ray at radius r with zenith cosine mu:
// t^2 + 2 b t - c = 0, b = r mu, c = R'^2 - r^2 (passed pre-factored), in
// the conjugate forms that do not cancel.  disc = b^2 + c.
__device__ __forceinline__ void sphere_roots(float b, float c, float& near_t, float& far_t, float& disc) {
    disc = b * b + c;
    float s = sqrtf(fmaxf(disc, 0.0f));
    if (b >= 0.0f) {
        near_t = -b - s;
        far_t = (b + s) > 0.0f ? c / (b + s) : -b + s;
    } else {
        far_t = -b + s;
        near_t = (s - b) > 0.0f ? -c / (s - b) : -b - s;
    }
}

// Trilinear CLAMP_TO_EDGE fetch of an RGBA32F 3D LUT at texel-space coordinates.
__device__ inline F3 lut3d_trilinear(const float4* __restrict__ lut, int w, int h, int d, float fx, float fy,
                                     float fz) {
    fx = fminf(fmaxf(fx, 0.0f), float(w - 1));
    fy = fminf(fmaxf(fy, 0.0f), float(h - 1));
    fz = fminf(fmaxf(fz, 0.0f), float(d - 1));
    int x0 = int(fx), y0 = int(fy), z0 = int(fz);
    int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1), z1 = min(z0 + 1, d - 1);
    float ax = fx - float(x0), ay = fy - float(y0), az = fz - float(z0);
    F3 r = f3(0.0f);
    for (int k = 0; k < 8; ++k) {
        int xx = (k & 1) ? x1 : x0, yy = (k & 2) ? y1 : y0, zz = (k & 4) ? z1 : z0;
        float wgt = ((k & 1) ? ax : 1.0f - ax) * ((k & 2) ? ay : 1.0f - ay) * ((k & 4) ? az : 1.0f - az);
        float4 v = lut[(zz * h + yy) * w + xx];
        r += F3{v.x, v.y, v.z} * wgt;
    }
    return r;
}

// SampleRayScatterVisibility (VolumetricCloudShadowInterface.glsl:10-13): the R16 UNORM froxel
// (LINEAR, CLAMP_TO_EDGE: GetShadowFroxel, VolumetricCloud.h:71-73) at (uv, dist / max distance),
// faded towards 1 beyond the froxel's far end.
__device__ inline float ray_scatter_visibility(const uint16_t* __restrict__ fr, int fw, int fh, int fd, float u,
                                               float v, float dist, float inv_max_dist) {
    const float w = dist * inv_max_dist;
    const float fx = fminf(fmaxf(u * float(fw) - 0.5f, 0.0f), float(fw - 1));
    const float fy = fminf(fmaxf(v * float(fh) - 0.5f, 0.0f), float(fh - 1));
    const float fz = fminf(fmaxf(w * float(fd) - 0.5f, 0.0f), float(fd - 1));
    const int x0 = int(fx), y0 = int(fy), z0 = int(fz);
    const int x1 = min(x0 + 1, fw - 1), y1 = min(y0 + 1, fh - 1), z1 = min(z0 + 1, fd - 1);
    const float ax = fx - float(x0), ay = fy - float(y0), az = fz - float(z0);
    float tex = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int xx = (k & 1) ? x1 : x0, yy = (k & 2) ? y1 : y0, zz = (k & 4) ? z1 : z0;
        const float wgt = ((k & 1) ? ax : 1.0f - ax) * ((k & 2) ? ay : 1.0f - ay) * ((k & 4) ? az : 1.0f - az);
        tex += float(fr[(zz * fh + yy) * fw + xx]) * (1.0f / 65535.0f) * wgt;
    }
    const float t = fminf(fmaxf(1.0f / w, 0.0f), 1.0f);
    return (1.0f - t) + tex * t;  // mix(1.0, tex, t)
}

// FP64 unprojection of (u, v, depth) through uInvMVP, relative to the camera.
__device__ inline F3 unproject(const float* m, float x, float y, float z, F3 cam, float* frag_dist) {
    double v[4] = {x, y, z, 1.0}, o[4];
    for (int row = 0; row < 4; ++row)
        o[row] = double(m[row]) * v[0] + double(m[4 + row]) * v[1] + double(m[8 + row]) * v[2] +
                 double(m[12 + row]) * v[3];
    double px = o[0] / o[3] - double(cam.x), py = o[1] / o[3] - double(cam.y), pz = o[2] / o[3] - double(cam.z);
    double len = sqrt(px * px + py * py + pz * pz);
    *frag_dist = float(len);
    return {float(px / len), float(py / len), float(pz / len)};
}

}  // namespace skyhip
