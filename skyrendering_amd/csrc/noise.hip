// Cloud-texture generators for gfx950 (NoiseGen.comp + shaders/Base/Noise.glsl):
// the 2D cloud map (RG8: Perlin-FBM coverage, Worley-FBM type), the 3D detail
// volume (R8 Perlin-Worley), the 3D Worley detail volume (R8 Worley-FBM) and the 2D
// displacement map (RGBA8 Perlin-FBM),
// plus the box-filter mip chain builder.  One lane per texel.  Hashes are
// exact uint32 arithmetic; FBM sums run in FP32 as in the shader.
#include <hip/hip_fp16.h>

#include "common.h"

namespace skyhip {

__device__ __forceinline__ uint32_t wang_hash(uint32_t seed) {
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}

// Gradient table of Noise.glsl:17-22, packed as 2-bit signed triples.
__device__ __forceinline__ float perlin_grad_dot(uint32_t h, float x, float y, float z) {
    constexpr int8_t g[16][3] = {{1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1},
                                 {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1},  {0, 1, -1}, {0, -1, -1},
                                 {1, 1, 0},  {-1, 1, 0},  {0, -1, 1}, {0, -1, -1}};
    int i = int(h & 15u);
    return float(g[i][0]) * x + float(g[i][1]) * y + float(g[i][2]) * z;
}

__device__ __forceinline__ float gmixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// PerlinNoise, Noise.glsl:28-60 (p >= 0 here: texture coordinates in [0, 1)).
__device__ inline float perlin_noise(float px, float py, float pz, uint32_t freq, uint32_t seed) {
    px *= float(freq);
    py *= float(freq);
    pz *= float(freq);
    float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    uint32_t i0 = uint32_t(int(fx)) % freq, j0 = uint32_t(int(fy)) % freq, k0 = uint32_t(int(fz)) % freq;
    uint32_t i1 = uint32_t(int(ceilf(px))) % freq, j1 = uint32_t(int(ceilf(py))) % freq,
             k1 = uint32_t(int(ceilf(pz))) % freq;
    float tx = px - fx, ty = py - fy, tz = pz - fz;
    float u = tx * tx * tx * (tx * (tx * 6.0f - 15.0f) + 10.0f);
    float v = ty * ty * ty * (ty * (ty * 6.0f - 15.0f) + 10.0f);
    float w = tz * tz * tz * (tz * (tz * 6.0f - 15.0f) + 10.0f);
    // GetPerlinGradients(i, j, k) = table[WangHash(seed + WangHash(i + WangHash(j + WangHash(k))))]
    uint32_t hk0 = wang_hash(k0), hk1 = wang_hash(k1);
    uint32_t hj0k0 = wang_hash(j0 + hk0), hj1k0 = wang_hash(j1 + hk0), hj0k1 = wang_hash(j0 + hk1),
             hj1k1 = wang_hash(j1 + hk1);
    auto G = [&](uint32_t i, uint32_t hjk) { return wang_hash(seed + wang_hash(i + hjk)); };
    float x1 = tx - 1.0f, y1 = ty - 1.0f, z1 = tz - 1.0f;
    float n000 = perlin_grad_dot(G(i0, hj0k0), tx, ty, tz), n100 = perlin_grad_dot(G(i1, hj0k0), x1, ty, tz);
    float n010 = perlin_grad_dot(G(i0, hj1k0), tx, y1, tz), n110 = perlin_grad_dot(G(i1, hj1k0), x1, y1, tz);
    float n001 = perlin_grad_dot(G(i0, hj0k1), tx, ty, z1), n101 = perlin_grad_dot(G(i1, hj0k1), x1, ty, z1);
    float n011 = perlin_grad_dot(G(i0, hj1k1), tx, y1, z1), n111 = perlin_grad_dot(G(i1, hj1k1), x1, y1, z1);
    return gmixf(gmixf(gmixf(n000, n100, u), gmixf(n010, n110, u), v),
                 gmixf(gmixf(n001, n101, u), gmixf(n011, n111, u), v), w);
}

// WorleyNoise (3D), Noise.glsl:81-101.
__device__ inline float worley_noise(float px, float py, float pz, uint32_t freq, uint32_t seed) {
    px *= float(freq);
    py *= float(freq);
    pz *= float(freq);
    uint32_t ix = uint32_t(floorf(px)), iy = uint32_t(floorf(py)), iz = uint32_t(floorf(pz));
    px += float(freq);
    py += float(freq);
    pz += float(freq);
    float min_d2 = 1e20f;
    for (uint32_t dk = 0; dk < 3; ++dk) {
        uint32_t gz = iz + freq - 1 + dk, sz = gz % freq;
        uint32_t hz0 = wang_hash(sz), hz1 = wang_hash(sz + 1), hz2 = wang_hash(sz + 2);
        for (uint32_t dj = 0; dj < 3; ++dj) {
            uint32_t gy = iy + freq - 1 + dj, sy = gy % freq;
            uint32_t hy0 = wang_hash(sy + hz0), hy1 = wang_hash(sy + hz1), hy2 = wang_hash(sy + hz2);
            for (uint32_t di = 0; di < 3; ++di) {
                uint32_t gx = ix + freq - 1 + di, sx = gx % freq;
                uint32_t r0 = wang_hash(seed + wang_hash(sx + hy0));
                uint32_t r1 = wang_hash(seed + wang_hash(sx + hy1));
                uint32_t r2 = wang_hash(seed + wang_hash(sx + hy2));
                float dx = px - (float(gx) + float(r0) * 2.3283064365386963e-10f);
                float dy = py - (float(gy) + float(r1) * 2.3283064365386963e-10f);
                float dz = pz - (float(gz) + float(r2) * 2.3283064365386963e-10f);
                min_d2 = fminf(min_d2, dx * dx + dy * dy + dz * dz);
            }
        }
    }
    return sqrtf(min_d2);
}

__device__ __forceinline__ float remap_to01(float x, float x0, float x1) {
    return fminf(fmaxf((x - x0) / (x1 - x0), 0.0f), 1.0f);
}

// PerlinFBM / WorleyFBM, NoiseGen.comp:20-50 (8 octaves).
__device__ inline float perlin_fbm(float x, float y, float z, SkyhipNoiseCreateInfo ci) {
    float res = 0.0f, a = 0.5f, sum_a = 0.0f;
    uint32_t f = ci.base_frequency;
    for (int c = 0; c < 8; ++c) {
        float n = perlin_noise(x, y, z, f, ci.seed) * 0.5f + 0.5f;
        res += remap_to01(n, ci.remap_min, ci.remap_max) * a;
        sum_a += a;
        f *= 2u;
        a *= 0.5f;
    }
    return res / sum_a;
}
__device__ inline float worley_fbm(float x, float y, float z, SkyhipNoiseCreateInfo ci) {
    float res = 0.0f, a = 0.5f, sum_a = 0.0f;
    uint32_t f = ci.base_frequency;
    for (int c = 0; c < 8; ++c) {
        float n = worley_noise(x, y, z, f, ci.seed);
        res += remap_to01(n, ci.remap_min, ci.remap_max) * a;
        sum_a += a;
        f *= 2u;
        a *= 0.5f;
    }
    return res / sum_a;
}

__device__ __forceinline__ uint8_t to_unorm8(float f) {
    return uint8_t(floorf(fminf(fmaxf(f, 0.0f), 1.0f) * 255.0f + 0.5f));
}

__global__ __launch_bounds__(256) void cloud_map_kernel(SkyhipNoiseCreateInfo density, SkyhipNoiseCreateInfo height,
                                                        uint8_t* __restrict__ out, int size) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= size * size) return;
    float x = (float(t % size) + 0.5f) / float(size), y = (float(t / size) + 0.5f) / float(size);
    uchar2 v;
    v.x = to_unorm8(perlin_fbm(x, y, 0.0f, density));
    v.y = to_unorm8(worley_fbm(x, y, 0.0f, height));
    reinterpret_cast<uchar2*>(out)[t] = v;
}

__global__ __launch_bounds__(256) void detail_kernel(SkyhipNoiseCreateInfo perlin, SkyhipNoiseCreateInfo worley,
                                                     uint8_t* __restrict__ out, int size) {
    long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)size * size * size) return;
    int ix = int(t % size), iy = int((t / size) % size), iz = int(t / ((long long)size * size));
    float x = (float(ix) + 0.5f) / float(size), y = (float(iy) + 0.5f) / float(size),
          z = (float(iz) + 0.5f) / float(size);
    float p = perlin_fbm(x, y, z, perlin);
    float w = worley_fbm(x, y, z, worley);
    out[t] = to_unorm8(fminf(fmaxf(w + p * (1.0f - w), 0.0f), 1.0f));  // RemapFrom01(p, w, 1)
}

// The Worley detail volume: WorleyFBM alone at texel centres, UNORM8.
__global__ __launch_bounds__(256) void worley_kernel(SkyhipNoiseCreateInfo worley, uint8_t* __restrict__ out,
                                                     int size) {
    long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)size * size * size) return;
    int ix = int(t % size), iy = int((t / size) % size), iz = int(t / ((long long)size * size));
    float x = (float(ix) + 0.5f) / float(size), y = (float(iy) + 0.5f) / float(size),
          z = (float(iz) + 0.5f) / float(size);
    out[t] = to_unorm8(worley_fbm(x, y, z, worley));
}

__global__ __launch_bounds__(256) void displacement_kernel(SkyhipNoiseCreateInfo perlin, uint8_t* __restrict__ out,
                                                           int size) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= size * size) return;
    float x = (float(t % size) + 0.5f) / float(size), y = (float(t / size) + 0.5f) / float(size);
    uchar4 v;
    uint8_t* c = &v.x;
    for (int i = 0; i < 4; ++i) {
        SkyhipNoiseCreateInfo ci = perlin;
        ci.seed += uint32_t(i);
        c[i] = to_unorm8(perlin_fbm(x, y, 0.0f, ci));
    }
    reinterpret_cast<uchar4*>(out)[t] = v;
}

// One mip step of a UNORM8 texture: box over 2x2(x2) source texels (indices
// clamped to the source level), (sum + n/2) / n.
__global__ __launch_bounds__(256) void downsample_kernel(const uint8_t* __restrict__ src, int sx, int sy, int sz,
                                                         uint8_t* __restrict__ dst, int dx, int dy, int dz, int dims,
                                                         int channels) {
    long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    long long total = (long long)dx * dy * dz * channels;
    if (t >= total) return;
    int c = int(t % channels);
    long long texel = t / channels;
    int x = int(texel % dx), y = int((texel / dx) % dy), z = int(texel / ((long long)dx * dy));
    int n = dims == 3 ? 8 : 4, sum = 0;
    for (int k = 0; k < n; ++k) {
        int xx = min(2 * x + (k & 1), sx - 1), yy = min(2 * y + ((k >> 1) & 1), sy - 1);
        int zz = dims == 3 ? min(2 * z + ((k >> 2) & 1), sz - 1) : 0;
        sum += src[(((long long)zz * sy + yy) * sx + xx) * channels + c];
    }
    dst[t] = uint8_t((sum + n / 2) / n);
}

// The path-tracing workload's voxel-cloud density (see SkyhipVoxelGridCreateInfo):
// fBm of PerlinNoise (unremapped), lognormal transform, ellipsoidal falloff,
// threshold, FP16 round-to-nearest-even.  One lane per texel.
__global__ __launch_bounds__(256) void voxel_grid_kernel(SkyhipVoxelGridCreateInfo ci, uint16_t* __restrict__ out,
                                                         int nx, int ny, int nz) {
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)nx * ny * nz) return;
    int x = int(i % nx), y = int((i / nx) % ny), z = int(i / ((long long)nx * ny));
    float p[3] = {(float(x) + 0.5f) / float(nx), (float(y) + 0.5f) / float(ny), (float(z) + 0.5f) / float(nz)};
    float res = 0.0f, a = 1.0f, sum_a = 0.0f;
    uint32_t f = ci.base_frequency;
    for (int o = 0; o < ci.octaves; ++o) {
        res += perlin_noise(p[0], p[1], p[2], f, ci.seed) * a;
        sum_a += a;
        f *= 2u;
        a *= 0.5f;
    }
    float fbm = res / sum_a;
    float e = 0.0f;
    for (int k = 0; k < 3; ++k) {
        float q = (p[k] - ci.ellipsoid_center[k]) / ci.ellipsoid_radii[k];
        e += q * q;
    }
    float falloff = fmaxf(1.0f - e, 0.0f);
    float density = fmaxf(expf(ci.lognormal_sigma * ci.fbm_inv_std * fbm) * falloff - ci.threshold, 0.0f);
    out[i] = __builtin_bit_cast(uint16_t, __float2half_rn(density));
}

int validate_noise(const SkyhipNoiseCreateInfo* ci) {
    if (ci == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "noise create info is NULL");
    if (ci->base_frequency == 0 || ci->base_frequency > (1u << 20))
        return fail(SKYHIP_ERR_INVALID_SIZE, "base_frequency %u outside [1, 2^20]", ci->base_frequency);
    if (!std::isfinite(ci->remap_min) || !std::isfinite(ci->remap_max))
        return fail(SKYHIP_ERR_NON_FINITE, "non-finite remap range");
    return SKYHIP_OK;
}

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_mip_levels(int32_t width, int32_t height, int32_t depth) {
    int m = width > height ? width : height;
    m = m > depth ? m : depth;
    if (m <= 0) return 0;
    int levels = 1;
    while ((m >> levels) > 0) ++levels;  // floor(log2(max)) + 1, ImageLoader.h:8-14
    return levels;
}

extern "C" size_t skyhip_mip_chain_bytes(int32_t width, int32_t height, int32_t depth, int32_t channels,
                                         int32_t levels) {
    size_t total = 0;
    for (int l = 0; l < levels; ++l) {
        size_t w = size_t(width >> l > 0 ? width >> l : 1), h = size_t(height >> l > 0 ? height >> l : 1),
               d = size_t(depth >> l > 0 ? depth >> l : 1);
        total += w * h * d * size_t(channels);
    }
    return total;
}

extern "C" int skyhip_noise_cloud_map(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* density,
                                      const SkyhipNoiseCreateInfo* height, uint8_t* out_rg8, int32_t size) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_noise(density));
    SKYHIP_CHECK(validate_noise(height));
    if (out_rg8 == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "output is NULL");
    if (size < 1 || size > 16384) return fail(SKYHIP_ERR_INVALID_SIZE, "cloud map size %d outside [1, 16384]", size);
    hipLaunchKernelGGL(cloud_map_kernel, dim3(div_up((long long)size * size, 256)), dim3(256), 0, ctx->stream,
                       *density, *height, out_rg8, size);
    SKYHIP_LAUNCH_CHECK("cloud_map_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_noise_detail(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* perlin,
                                   const SkyhipNoiseCreateInfo* worley, uint8_t* out_r8, int32_t size) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_noise(perlin));
    SKYHIP_CHECK(validate_noise(worley));
    if (out_r8 == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "output is NULL");
    if (size < 1 || size > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "detail size %d outside [1, 1024]", size);
    hipLaunchKernelGGL(detail_kernel, dim3(div_up((long long)size * size * size, 256)), dim3(256), 0, ctx->stream,
                       *perlin, *worley, out_r8, size);
    SKYHIP_LAUNCH_CHECK("detail_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_noise_displacement(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* perlin, uint8_t* out_rgba8,
                                         int32_t size) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_noise(perlin));
    if (out_rgba8 == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "output is NULL");
    if (size < 1 || size > 16384) return fail(SKYHIP_ERR_INVALID_SIZE, "displacement size %d outside [1, 16384]", size);
    hipLaunchKernelGGL(displacement_kernel, dim3(div_up((long long)size * size, 256)), dim3(256), 0, ctx->stream,
                       *perlin, out_rgba8, size);
    SKYHIP_LAUNCH_CHECK("displacement_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_noise_worley(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* worley, uint8_t* out_r8,
                                   int32_t size) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_noise(worley));
    if (out_r8 == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "output is NULL");
    if (size < 1 || size > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "Worley volume size %d outside [1, 1024]", size);
    hipLaunchKernelGGL(worley_kernel, dim3(div_up((long long)size * size * size, 256)), dim3(256), 0, ctx->stream,
                       *worley, out_r8, size);
    SKYHIP_LAUNCH_CHECK("worley_kernel");
    return SKYHIP_OK;
}

// Cell-corner copies of a REPEAT texture's level 0 (size s a power of two, indices mod s):
// a 2D RG8 cell (x, y) packs texels (x, y), (x + 1, y) in .x and (x, y + 1), (x + 1, y + 1) in
// .y, two bytes each; a 3D R8 cell (x, y, z) packs the eight texels (x + a, y + b, z + c) as
// byte a + 2b + 4c.
__global__ __launch_bounds__(256) void cells_2d_rg8_kernel(const uint16_t* __restrict__ t, int ls,
                                                           uint2* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, s = 1u << ls, m = s - 1u;
    if (i >= s * s) return;
    const uint32_t x = i & m, y = i >> ls, x1 = (x + 1u) & m, y1 = (y + 1u) & m;
    const uint32_t r0 = y << ls, r1 = y1 << ls;
    out[i] = make_uint2(uint32_t(t[r0 | x]) | (uint32_t(t[r0 | x1]) << 16), uint32_t(t[r1 | x]) | (uint32_t(t[r1 | x1]) << 16));
}

__global__ __launch_bounds__(256) void cells_3d_r8_kernel(const uint8_t* __restrict__ t, int ls,
                                                          uint2* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x, s = 1u << ls, m = s - 1u;
    if (i >= s * s * s) return;
    const uint32_t x = i & m, y = (i >> ls) & m, z = i >> (2 * ls);
    uint32_t w[2] = {0u, 0u};
    for (int k = 0; k < 8; ++k) {
        const uint32_t xx = (x + uint32_t(k & 1)) & m, yy = (y + uint32_t((k >> 1) & 1)) & m,
                       zz = (z + uint32_t(k >> 2)) & m;
        w[k >> 2] |= uint32_t(t[(zz << (2 * ls)) | (yy << ls) | xx]) << (8 * (k & 3));
    }
    out[i] = make_uint2(w[0], w[1]);
}

static int cells_args(SkyhipContext* ctx, const uint8_t* level0, int32_t size, uint8_t* out, int max_size,
                      int align_in, int& ls) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!level0 || !out) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (size < 2 || size > max_size || (size & (size - 1)) != 0)
        return fail(SKYHIP_ERR_INVALID_SIZE, "texture size %d is not a power of two in [2, %d]", size, max_size);
    if (reinterpret_cast<uintptr_t>(out) % 8 != 0 || reinterpret_cast<uintptr_t>(level0) % align_in != 0)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "misaligned texture or cell buffer");
    ls = 0;
    while ((1 << ls) < size) ++ls;
    return SKYHIP_OK;
}

extern "C" int skyhip_texture_cells_2d_rg8(SkyhipContext* ctx, const uint8_t* level0, int32_t size,
                                           uint8_t* out_cells) {
    int ls;
    SKYHIP_CHECK(cells_args(ctx, level0, size, out_cells, 16384, 2, ls));
    hipLaunchKernelGGL(cells_2d_rg8_kernel, dim3(div_up((long long)size * size, 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const uint16_t*>(level0), ls, reinterpret_cast<uint2*>(out_cells));
    SKYHIP_LAUNCH_CHECK("cells_2d_rg8_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_texture_cells_3d_r8(SkyhipContext* ctx, const uint8_t* level0, int32_t size,
                                          uint8_t* out_cells) {
    int ls;
    SKYHIP_CHECK(cells_args(ctx, level0, size, out_cells, 1024, 1, ls));
    hipLaunchKernelGGL(cells_3d_r8_kernel, dim3(div_up((long long)size * size * size, 256)), dim3(256), 0,
                       ctx->stream, level0, ls, reinterpret_cast<uint2*>(out_cells));
    SKYHIP_LAUNCH_CHECK("cells_3d_r8_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_generate_mipmaps(SkyhipContext* ctx, uint8_t* texture, int32_t dims, int32_t channels,
                                       int32_t width, int32_t height, int32_t depth, int32_t levels) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (texture == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "texture is NULL");
    if ((dims != 2 && dims != 3) || channels < 1 || channels > 4)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "dims must be 2 or 3 and channels 1..4");
    if (dims == 2) depth = 1;
    if (width < 1 || height < 1 || depth < 1 || levels < 1 || levels > skyhip_mip_levels(width, height, depth))
        return fail(SKYHIP_ERR_INVALID_SIZE, "bad texture size or level count");
    size_t off = 0;
    int sx = width, sy = height, sz = depth;
    for (int l = 1; l < levels; ++l) {
        size_t src_bytes = size_t(sx) * sy * sz * channels;
        int dx = sx > 1 ? sx / 2 : 1, dy = sy > 1 ? sy / 2 : 1, dz = dims == 3 ? (sz > 1 ? sz / 2 : 1) : 1;
        hipLaunchKernelGGL(downsample_kernel, dim3(div_up((long long)dx * dy * dz * channels, 256)), dim3(256), 0,
                           ctx->stream, texture + off, sx, sy, sz, texture + off + src_bytes, dx, dy, dz, dims,
                           channels);
        SKYHIP_LAUNCH_CHECK("downsample_kernel");
        off += src_bytes;
        sx = dx, sy = dy, sz = dz;
    }
    return SKYHIP_OK;
}

extern "C" int skyhip_voxel_grid_generate(SkyhipContext* ctx, const SkyhipVoxelGridCreateInfo* info,
                                          uint16_t* out_grid, int32_t dim_x, int32_t dim_y, int32_t dim_z) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (info == nullptr || out_grid == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dim_x < 8 || dim_y < 8 || dim_z < 8 || dim_x > 1024 || dim_y > 1024 || dim_z > 1024)
        return fail(SKYHIP_ERR_INVALID_SIZE, "grid dims outside [8, 1024]");
    if (info->base_frequency < 1 || info->base_frequency > (1u << 16) || info->octaves < 1 || info->octaves > 12)
        return fail(SKYHIP_ERR_INVALID_SIZE, "base_frequency outside [1, 65536] or octaves outside [1, 12]");
    if (!finite_block(&info->fbm_inv_std, 9)) return fail(SKYHIP_ERR_NON_FINITE, "non-finite grid parameter");
    for (int k = 0; k < 3; ++k)
        if (!(info->ellipsoid_radii[k] > 0.0f)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "ellipsoid radii must be > 0");
    long long n = (long long)dim_x * dim_y * dim_z;
    hipLaunchKernelGGL(voxel_grid_kernel, dim3(unsigned(div_up(n, 256))), dim3(256), 0, ctx->stream, *info, out_grid,
                       dim_x, dim_y, dim_z);
    SKYHIP_LAUNCH_CHECK("voxel_grid_kernel");
    return SKYHIP_OK;
}
