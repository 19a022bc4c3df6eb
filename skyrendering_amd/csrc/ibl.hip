// The environment-luminance cube of the atmosphere renderer and its image-based-lighting
// precomputation:
//  - env_luminance_kernel: one thread per cube texel, the sky-view LUT pair looked up along
//    the texel's direction plus the lit ground (AtmosphereRenderer.glsl:248-275), stored as
//    RGBA16F (a = 0);
//  - cube_mip_kernel, cube_mip_tail_kernel: the cube's mip chain, a 2x2 box per level
//    (glGenerateTextureMipmap, AtmosphereRenderer.cpp:242; the filter is the implementation's,
//    here a box), the levels of at most 64^2 texels a face in one launch;
//  - prefilter_kernel: the 9 spherical-harmonic radiance coefficients (EnvRadianceSH.comp), one
//    workgroup per coefficient over 1024 stratified directions summed by the shader's halving
//    tree, and in the same launch the GGX-prefiltered radiance cube, one level per roughness
//    (PrefilterRadiance.comp, IBL.cpp:25-45), a texel's samples spread over a group of lanes.
// Each pass is tiny (at most 98,304 texels); the sequence is bound by launch latency.
#include "sky_env.h"

namespace skyhip {

int validate_atmosphere(const SkyhipAtmosphereBufferData* b);
int validate_render(const SkyhipAtmosphereRenderBufferData* b);
int validate_lut(const void* p, int w, int h, const char* what);

namespace {

__device__ inline ushort4 to_rgba16f(F3 c, float a) {
    ushort4 o;
    o.x = __half_as_ushort(__float2half_rn(c.x));
    o.y = __half_as_ushort(__float2half_rn(c.y));
    o.z = __half_as_ushort(__float2half_rn(c.z));
    o.w = __half_as_ushort(__float2half_rn(a));
    return o;
}

__device__ inline float h2f(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }

bool is_pow2(int w) { return w > 0 && (w & (w - 1)) == 0; }
int log2i(int w) {
    int l = 0;
    while ((1 << (l + 1)) <= w) ++l;
    return l;
}

}  // namespace

__global__ __launch_bounds__(256) void env_luminance_kernel(AtmDev A, EnvFrame e, const float4* __restrict__ T,
                                                            int tw, int th, const float4* __restrict__ sv_l,
                                                            const float4* __restrict__ sv_t, int sw, int sh,
                                                            ushort4* __restrict__ out, int w) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 6 * w * w) return;
    const int face = i / (w * w), y = (i / w) % w, x = i % w;
    const F3 d = cube_texel_dir(face, x, y, w);
    out[i] = to_rgba16f(env_luminance(A, e, T, tw, th, sv_l, sv_t, sw, sh, d), 0.0f);
}

// Level `level` of the cube from level - 1: the mean of the 2x2 parent texels per channel.
__global__ __launch_bounds__(256) void cube_mip_kernel(ushort4* cube, int w, int level) {
    const int wl = max(w >> level, 1), wp = max(w >> (level - 1), 1);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 6 * wl * wl) return;
    const int face = i / (wl * wl), y = (i / wl) % wl, x = i % wl;
    const ushort4* src = cube + cube_level_offset(w, level - 1) + (long long)face * wp * wp;
    const int x0 = min(2 * x, wp - 1), x1 = min(2 * x + 1, wp - 1), y0 = min(2 * y, wp - 1), y1 = min(2 * y + 1, wp - 1);
    const ushort4 a = src[y0 * wp + x0], b = src[y0 * wp + x1], c = src[y1 * wp + x0], d = src[y1 * wp + x1];
    ushort4 o;
    o.x = __half_as_ushort(__float2half_rn(((h2f(a.x) + h2f(b.x)) + (h2f(c.x) + h2f(d.x))) * 0.25f));
    o.y = __half_as_ushort(__float2half_rn(((h2f(a.y) + h2f(b.y)) + (h2f(c.y) + h2f(d.y))) * 0.25f));
    o.z = __half_as_ushort(__float2half_rn(((h2f(a.z) + h2f(b.z)) + (h2f(c.z) + h2f(d.z))) * 0.25f));
    o.w = __half_as_ushort(__float2half_rn(((h2f(a.w) + h2f(b.w)) + (h2f(c.w) + h2f(d.w))) * 0.25f));
    cube[cube_level_offset(w, level) + (long long)face * wl * wl + y * wl + x] = o;
}

// The mean of a 2x2 parent block per channel (cube_mip_kernel's arithmetic).
__device__ inline ushort4 box4(ushort4 a, ushort4 b, ushort4 c, ushort4 d) {
    ushort4 o;
    o.x = __half_as_ushort(__float2half_rn(((h2f(a.x) + h2f(b.x)) + (h2f(c.x) + h2f(d.x))) * 0.25f));
    o.y = __half_as_ushort(__float2half_rn(((h2f(a.y) + h2f(b.y)) + (h2f(c.y) + h2f(d.y))) * 0.25f));
    o.z = __half_as_ushort(__float2half_rn(((h2f(a.z) + h2f(b.z)) + (h2f(c.z) + h2f(d.z))) * 0.25f));
    o.w = __half_as_ushort(__float2half_rn(((h2f(a.w) + h2f(b.w)) + (h2f(c.w) + h2f(d.w))) * 0.25f));
    return o;
}

// Levels first_level .. levels - 1 of one face (blockIdx.x) in one workgroup, for the levels of at
// most kMipTailWidth^2 texels a face: the first from the parent level in global memory, each
// later one from the previous level kept in LDS, every level also stored to the cube.  The same
// arithmetic as cube_mip_kernel level by level, without a launch per level.
constexpr int kMipTailWidth = 64;
__global__ __launch_bounds__(1024) void cube_mip_tail_kernel(ushort4* cube, int w, int first_level, int levels) {
    __shared__ ushort4 s_a[kMipTailWidth * kMipTailWidth];
    __shared__ ushort4 s_b[(kMipTailWidth / 2) * (kMipTailWidth / 2)];
    const int face = blockIdx.x;
    for (int l = first_level; l < levels; ++l) {
        const int wl = max(w >> l, 1), wp = max(w >> (l - 1), 1);
        const bool from_global = l == first_level;
        const ushort4* gsrc = cube + cube_level_offset(w, l - 1) + (long long)face * wp * wp;
        const ushort4* lsrc = ((l - first_level) & 1) ? s_a : s_b;  // the previous level's buffer
        ushort4* ldst = ((l - first_level) & 1) ? s_b : s_a;
        ushort4* gdst = cube + cube_level_offset(w, l) + (long long)face * wl * wl;
        for (int t = threadIdx.x; t < wl * wl; t += 1024) {
            const int y = t / wl, x = t % wl;
            const int x0 = min(2 * x, wp - 1), x1 = min(2 * x + 1, wp - 1), y0 = min(2 * y, wp - 1),
                      y1 = min(2 * y + 1, wp - 1);
            const ushort4* src = from_global ? gsrc : lsrc;
            const ushort4 o = box4(src[y0 * wp + x0], src[y0 * wp + x1], src[y1 * wp + x0], src[y1 * wp + x1]);
            ldst[t] = o;
            gdst[t] = o;
        }
        __syncthreads();
    }
}

// EnvRadianceSH.comp: coefficient `index` from 1024 directions on a 32 x 32 stratified grid
// (uniform in cos theta and phi, y up), each sample weighted 4 pi / 1024, summed by the shader's
// halving tree over the 1024 values (at half = 512, 256, ..., 1: s[i] += s[i + half]).  A
// 256-thread workgroup holds samples t, t + 256, t + 512 and t + 768 in thread t and forms the
// tree's first two levels in registers, (s_t + s_(t+512)) + (s_(t+256) + s_(t+768)), the rest
// over LDS: the same additions in the same order.
__device__ inline F3 sh9_sample(const ushort4* __restrict__ cube, int w, int levels, int index, int li) {
    const float unit_theta = (0.5f + float(li >> 5)) / 32.0f;
    const float unit_phi = (0.5f + float(li & 31)) / 32.0f;
    const float cos_theta = 1.0f - 2.0f * unit_theta;
    const float sin_theta = sqrtf(clamp01(1.0f - cos_theta * cos_theta));
    const float phi = 2.0f * kPiF * unit_phi;
    const F3 dir = {cosf(phi) * sin_theta, cos_theta, sinf(phi) * sin_theta};
    const F3 radiance = cube_sample_lod(cube, w, levels, dir, 0.0f);
    const float Y00 = 0.282095f, Y1n = 0.488603f, Y2n = 1.092548f, Y20 = 0.315392f, Y22 = 0.546274f;
    float c;
    switch (index) {
        case 0: c = Y00; break;
        case 1: c = Y1n * dir.y; break;
        case 2: c = Y1n * dir.z; break;
        case 3: c = Y1n * dir.x; break;
        case 4: c = Y2n * dir.x * dir.y; break;
        case 5: c = Y2n * dir.y * dir.z; break;
        case 6: c = Y20 * (3.0f * dir.z * dir.z - 1.0f); break;
        case 7: c = Y2n * dir.x * dir.z; break;
        default: c = Y22 * (dir.x * dir.x - dir.y * dir.y); break;
    }
    return radiance * (c * (4.0f * kPiF / 1024.0f));
}
__device__ inline void sh9_block(const ushort4* __restrict__ cube, int w, int levels, int index,
                                 float4* __restrict__ out, F3* s) {
    const int t = threadIdx.x;
    const F3 lo = sh9_sample(cube, w, levels, index, t) + sh9_sample(cube, w, levels, index, t + 512);
    const F3 hi = sh9_sample(cube, w, levels, index, t + 256) + sh9_sample(cube, w, levels, index, t + 768);
    s[t] = lo + hi;
    __syncthreads();
    for (int half = 128; half >= 1; half >>= 1) {
        if (t < half) s[t] = s[t] + s[t + half];
        __syncthreads();
    }
    if (t == 0) out[index] = make_float4(s[0].x, s[0].y, s[0].z, 0.0f);
}
constexpr int kSh9Blocks = 9;

__device__ inline uint32_t reverse_bits32(uint32_t b) {  // ReverseBits32 (shaders/Base/Noise.glsl:104-111)
    b = (b << 16) | (b >> 16);
    b = ((b & 0x00ff00ffu) << 8) | ((b & 0xff00ff00u) >> 8);
    b = ((b & 0x0f0f0f0fu) << 4) | ((b & 0xf0f0f0f0u) >> 4);
    b = ((b & 0x33333333u) << 2) | ((b & 0xccccccccu) >> 2);
    b = ((b & 0x55555555u) << 1) | ((b & 0xaaaaaaaau) >> 1);
    return b;
}

// CreateOrthonormalBasis (shaders/Base/Common.glsl:44-51, Duff et al. 2017)
__device__ inline void onb_frame(F3 N, F3& t0, F3& t1) {
    const float s = N.z >= 0.0f ? 1.0f : -1.0f;
    const float aa = -1.0f / (s + N.z);
    const float b = N.x * N.y * aa;
    t0 = {1.0f + s * N.x * N.x * aa, s * b, -s * N.x};
    t1 = {b, s + N.y * N.y * aa, -N.y};
}

constexpr int kMaxPrefilterLevels = 13;  // roughness_count <= log2(4096) + 1

// PrefilterRadiance.comp for level `level` (width wl) of the output cube at roughness r.  Each
// texel's samples are spread over a group of `group` = n lanes (the level's sample count, at most
// 64; 1 at roughness 0), 256 / n texels to a workgroup: lane k evaluates sample k into LDS, and
// the group's first lane adds the terms in sample order, so the sums are the
// one-thread-per-texel loop's.
// The rough levels have few texels (384 at 8^2 x 6) and up to 64 samples each, which one thread
// per texel left as a chain of dependent cube fetches.
// All levels run in one launch, after the kSh9Blocks workgroups that compute the SH9
// coefficients (sh9_block; both read the environment cube only): workgroups kSh9Blocks +
// [block_start[l], block_start[l + 1]) build level l.
struct PrefilterLevels {
    int count;
    int block_start[kMaxPrefilterLevels + 1];
    float roughness[kMaxPrefilterLevels];
    int group[kMaxPrefilterLevels];  // lanes per texel (the sample count)
};

__global__ __launch_bounds__(256) void prefilter_kernel(const ushort4* __restrict__ env, int ew, int elevels,
                                                        float4* __restrict__ sh9_out, ushort4* __restrict__ out,
                                                        int pw, PrefilterLevels lv) {
    __shared__ float4 s_term[256];
    __shared__ F3 s_sh[256];
    if (int(blockIdx.x) < kSh9Blocks) {
        sh9_block(env, ew, elevels, int(blockIdx.x), sh9_out, s_sh);
        return;
    }
    const int block = int(blockIdx.x) - kSh9Blocks;
    int level = 0;
    while (level + 1 < lv.count && block >= lv.block_start[level + 1]) ++level;
    const float roughness = lv.roughness[level];
    const int group = lv.group[level];
    const int wl = max(pw >> level, 1);
    const int per_block = 256 / group;
    const int j = int(threadIdx.x) / group, k = int(threadIdx.x) - j * group;
    const int i = (block - lv.block_start[level]) * per_block + j;
    const bool texel = j < per_block && i < 6 * wl * wl;
    const int face = i / (wl * wl), y = (i / wl) % wl, x = i % wl;
    const F3 R = cube_texel_dir(face, x, y, wl);
    const F3 N = R, V = R;
    const float a = roughness * roughness, a2 = a * a;
    // uint(mix(1, 64, pow(roughness, 0.3))), evaluated on the host with the group sizes
    const uint32_t n = uint32_t(group);
    const float inv_sa_texel = (6.0f * float(wl) * float(wl)) / (4.0f * kPiF);
    float4 term = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // (sample * NoL, NoL), NoL = 0: no contribution
    if (texel && uint32_t(k) < n) {
        F3 t0, t1;
        onb_frame(N, t0, t1);
        // Hammersley (shaders/Base/Noise.glsl:113-117) with no random offset
        const float q = float(uint32_t(k)) / float(n);
        const float e1 = q - floorf(q);
        const float e2 = float(reverse_bits32(uint32_t(k))) * 2.3283064365386963e-10f;
        // ImportanceSampleGGX (shaders/Base/BRDF.glsl:48-67)
        const float phi = 2.0f * kPiF * e1;
        const float cos_t = sqrtf((1.0f - e2) / (1.0f + (a2 - 1.0f) * e2));
        const float sin_t = sqrtf(1.0f - cos_t * cos_t);
        const F3 H = t0 * (sin_t * cosf(phi)) + t1 * (sin_t * sinf(phi)) + N * cos_t;
        const F3 L = H * (2.0f * dot(V, H)) - V;
        const float NoL = clamp01(dot(N, L));
        if (NoL > 0.0f) {
            const float NoH = clamp01(dot(N, H)), HoV = clamp01(dot(H, V));
            const float dg = (NoH * a2 - NoH) * NoH + 1.0f;  // D_GGX (BRDF.glsl:36-40)
            const float D = a2 / (kPiF * dg * dg);
            const float pdf = fmaxf(D * NoH / (4.0f * HoV), 0.0001f);
            const float sa_sample = 1.0f / fmaxf(float(n) * pdf, 0.00001f);
            const float mip = roughness == 0.0f ? 0.0f : 0.5f * log2f(sa_sample * inv_sa_texel) + 2.5f;
            const F3 c = cube_sample_lod(env, ew, elevels, L, mip) * NoL;
            term = make_float4(c.x, c.y, c.z, NoL);
        }
    }
    s_term[threadIdx.x] = term;
    __syncthreads();
    if (!texel || k != 0) return;
    F3 color = f3(0.0f);
    float total = 0.0f;
    for (uint32_t m = 0; m < n; ++m) {
        const float4 t = s_term[int(threadIdx.x) + int(m)];
        if (t.w > 0.0f) {
            color += F3{t.x, t.y, t.z};
            total += t.w;
        }
    }
    out[cube_level_offset(pw, level) + i] = to_rgba16f(color * (1.0f / total), 0.0f);
}

}  // namespace skyhip

using namespace skyhip;

extern "C" long long skyhip_cube_texels(int32_t width, int32_t levels) {
    if (width < 1 || levels < 0) return 0;
    return cube_level_offset(width, levels);
}

extern "C" int skyhip_environment_luminance(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                            const SkyhipAtmosphereRenderBufferData* render,
                                            const float* transmittance, int32_t transmittance_width,
                                            int32_t transmittance_height, const float* sky_view_luminance,
                                            const float* sky_view_transmittance, int32_t sky_view_width,
                                            int32_t sky_view_height, void* out_cube_rgba16f, int32_t width) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!atmosphere || !render || !out_cube_rgba16f) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    SKYHIP_CHECK(validate_lut(transmittance, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(sky_view_luminance, sky_view_width, sky_view_height, "sky-view luminance"));
    SKYHIP_CHECK(validate_lut(sky_view_transmittance, sky_view_width, sky_view_height, "sky-view transmittance"));
    if (!is_pow2(width) || width > 4096) return fail(SKYHIP_ERR_INVALID_SIZE, "cube width must be a power of two <= 4096");
    const AtmDev A = make_atm(*atmosphere);
    const EnvFrame e = make_env_frame(*render, *atmosphere);
    ushort4* cube = reinterpret_cast<ushort4*>(out_cube_rgba16f);
    hipLaunchKernelGGL(env_luminance_kernel, dim3(div_up(6LL * width * width, 256)), dim3(256), 0, ctx->stream, A, e,
                       reinterpret_cast<const float4*>(transmittance), transmittance_width, transmittance_height,
                       reinterpret_cast<const float4*>(sky_view_luminance),
                       reinterpret_cast<const float4*>(sky_view_transmittance), sky_view_width, sky_view_height, cube,
                       width);
    SKYHIP_LAUNCH_CHECK("env_luminance_kernel");
    const int levels = log2i(width) + 1;
    int l = 1;
    for (; l < levels && (width >> l) > kMipTailWidth; ++l) {  // levels wider than the tail's buffer
        const int wl = width >> l;
        hipLaunchKernelGGL(cube_mip_kernel, dim3(div_up(6LL * wl * wl, 256)), dim3(256), 0, ctx->stream, cube, width,
                           l);
    }
    SKYHIP_LAUNCH_CHECK("cube_mip_kernel");
    if (l < levels) {
        hipLaunchKernelGGL(cube_mip_tail_kernel, dim3(6), dim3(1024), 0, ctx->stream, cube, width, l, levels);
        SKYHIP_LAUNCH_CHECK("cube_mip_tail_kernel");
    }
    return SKYHIP_OK;
}

extern "C" int skyhip_ibl_precompute(SkyhipContext* ctx, const void* environment_cube_rgba16f,
                                     int32_t environment_width, float* out_sh9_rgba, void* out_prefiltered_rgba16f,
                                     int32_t prefiltered_width, int32_t roughness_count) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!environment_cube_rgba16f || !out_sh9_rgba || !out_prefiltered_rgba16f)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!is_pow2(environment_width) || environment_width > 4096)
        return fail(SKYHIP_ERR_INVALID_SIZE, "environment width must be a power of two <= 4096");
    if (!is_pow2(prefiltered_width) || prefiltered_width > 4096)
        return fail(SKYHIP_ERR_INVALID_SIZE, "prefiltered width must be a power of two <= 4096");
    if (roughness_count < 1 || roughness_count > kMaxPrefilterLevels || (1 << (roughness_count - 1)) > prefiltered_width)
        return fail(SKYHIP_ERR_INVALID_SIZE, "roughness_count outside [1, log2(prefiltered_width) + 1]");
    const ushort4* env = reinterpret_cast<const ushort4*>(environment_cube_rgba16f);
    const int elevels = log2i(environment_width) + 1;
    ushort4* out = reinterpret_cast<ushort4*>(out_prefiltered_rgba16f);
    PrefilterLevels lv{};
    lv.count = roughness_count;
    long long blocks = 0;
    for (int l = 0; l < roughness_count; ++l) {
        const int wl = prefiltered_width >> l;
        lv.roughness[l] = roughness_count > 1 ? float(l) / float(roughness_count - 1) : 0.0f;
        // The level's sample count uint(mix(1, 64, pow(roughness, 0.3))) in FP32 (1 at roughness 0,
        // 64 at 1; in between, for every level of up to 13, at least 0.011 from an integer, so no
        // powf rounding moves it), which the kernel uses as its group size and its sample count.
        const float p = std::pow(lv.roughness[l], 0.3f);
        lv.group[l] = std::min(std::max(int(uint32_t(1.0f * (1.0f - p) + 64.0f * p)), 1), 64);
        lv.block_start[l] = int(blocks);
        blocks += div_up(6LL * wl * wl, 256 / lv.group[l]);
    }
    lv.block_start[roughness_count] = int(blocks);
    hipLaunchKernelGGL(prefilter_kernel, dim3(unsigned(kSh9Blocks + blocks)), dim3(256), 0, ctx->stream, env,
                       environment_width, elevels, reinterpret_cast<float4*>(out_sh9_rgba), out, prefiltered_width, lv);
    SKYHIP_LAUNCH_CHECK("prefilter_kernel");
    return SKYHIP_OK;
}
