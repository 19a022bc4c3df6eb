// Atmosphere LUT kernels for gfx950: transmittance, multiple scattering,
// sky view and aerial perspective, plus their C-ABI entry points.
//
// Layout: every LUT is RGBA32F in GL texel order, one lane per texel (one
// wave64 per texel for multiple scattering, whose 64 sphere directions map to
// the 64 lanes), 256-lane workgroups, 1D grids over texels so stores coalesce.
#include "lut_texels.h"

#include <cmath>

namespace skyhip {

// ---------------------------------------------------------------------------
// Transmittance (Atmosphere.glsl:71-88, :311-337)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void transmittance_kernel(SkyhipAtmosphereBufferData b, float4* __restrict__ out,
                                                            int w, int h) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w * h) return;
    out[t] = transmittance_texel_rgba(b, t % w, t / w, w, h);
}

// ---------------------------------------------------------------------------
// Multiple scattering (Atmosphere.glsl:344-438): one wave64 per texel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void multiscattering_kernel(SkyhipAtmosphereBufferData b,
                                                              const float4* __restrict__ T, int tw, int th,
                                                              float4* __restrict__ out, int w, int h) {
    const AtmDev A = make_atm(b);
    const int texel = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (texel >= w * h) return;  // wave-uniform
    float4 v = multiscattering_texel(A, T, tw, th, texel % w, texel / w, w, h, lane);
    if (lane == 0) out[texel] = v;
}

// ---------------------------------------------------------------------------
// Sky view (AtmosphereRenderer.glsl:74-186)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sky_view_kernel(SkyhipAtmosphereBufferData b,
                                                       SkyhipAtmosphereRenderBufferData rbd,
                                                       const float4* __restrict__ T, int tw, int th,
                                                       const float4* __restrict__ MS, int mw, int mh,
                                                       const uint16_t* __restrict__ blue_noise, int bn_size,
                                                       float4* __restrict__ out_l, float4* __restrict__ out_t, int w,
                                                       int h) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w * h) return;
    const AtmDev A = make_atm(b);
    const RenderDev R = make_render(rbd);
    F3 L, Tr;
    sky_view_texel(A, R, T, tw, th, MS, mw, mh, blue_noise, bn_size, t % w, t / w, w, h, L, Tr);
    out_l[t] = make_float4(L.x, L.y, L.z, 0.0f);
    out_t[t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
}

// ---------------------------------------------------------------------------
// Aerial perspective (AtmosphereRenderer.glsl:191-243)
// ---------------------------------------------------------------------------
// kApLanes consecutive lanes share a froxel (scattered_luminance_split): a 32^3 LUT is 32,768
// froxels, and one lane each would leave half the SIMDs idle and the rest one wave deep.
constexpr int kApLanes = 4;

__global__ __launch_bounds__(256) void aerial_perspective_kernel(
    SkyhipAtmosphereBufferData b, SkyhipAtmosphereRenderBufferData rbd, const float4* __restrict__ T, int tw, int th,
    const float4* __restrict__ MS, int mw, int mh, const uint16_t* __restrict__ blue_noise, int bn_size,
    float4* __restrict__ out_l, float4* __restrict__ out_t, int w, int h, int d) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    const int n_texels = w * h * d;
    const int part = int(g % kApLanes);
    const bool valid = g / kApLanes < n_texels;
    const int t = valid ? int(g / kApLanes) : 0;  // every lane of a group reaches the shuffles
    const AtmDev A = make_atm(b);
    const RenderDev R = make_render(rbd);
    const AerialRay ray = aerial_ray(A, R, rbd.inv_view_projection, blue_noise, bn_size, t % w, (t / w) % h,
                                     t / (w * h), w, h, d);
    F3 L, Tr;
    scattered_luminance_split<kApLanes>(A, T, tw, th, MS, mw, mh, ray.s, ray.marching_distance, R.ap_steps,
                                        ray.start_i, ray.marching_distance > 0.0f, part, L, Tr);
    if (part == 0 && valid) {
        out_l[t] = make_float4(L.x, L.y, L.z, 0.0f);
        out_t[t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Latency-bound form of one frame's LUT set (skyhip_atmosphere_lut_set).  A single set has
// too few texels to fill 256 CUs one lane per texel (the 256x64 transmittance LUT is one wave
// per CU, the 192x108 sky view 1.3), so each texel's march is split over consecutive lanes
// and the parts combined in order, and the sky-view and aerial-perspective passes, which read
// only T and MS, run as one launch.  Same quantities as the per-pass kernels; the FP32 sums
// associate differently (parts add in order), which is why the batched builder, which must
// match the per-pass kernels bit for bit, does not use these.
// ---------------------------------------------------------------------------
constexpr int kSetTransmittanceParts = 4;
constexpr int kSetSkyViewParts = 3;
constexpr int kSetAerialParts = 8;

template <class S, int kParts>
__device__ inline F3 transmittance_texel_split(const SkyhipAtmosphereBufferData& b, int ix, int iy, int w, int h,
                                               int part) {
    const int n = transmittance_step_count(b);
    const int c = (n + kParts - 1) / kParts, first = min(part * c, n);
    S od[3], dx;
    transmittance_od<S>(b, ix, iy, w, h, first, max(0, min(c, n - first)), od, dx);
    S acc[3] = {od[0], od[1], od[2]};
    for (int p = 1; p < kParts; ++p)
        for (int k = 0; k < 3; ++k) acc[k] += __shfl_down(od[k], p, 64);
    return transmittance_from_od<S>(b, acc, dx);
}

__global__ __launch_bounds__(256) void set_transmittance_kernel(SkyhipAtmosphereBufferData b,
                                                                float4* __restrict__ out, int w, int h) {
    constexpr int P = kSetTransmittanceParts;
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int t = g / P, part = g % P;
    if (t >= w * h) return;  // whole groups of P lanes leave together
    const int ix = t % w, iy = t / w;
    const F3 v = ix >= transmittance_fp64_column(w) ? transmittance_texel_split<double, P>(b, ix, iy, w, h, part)
                                                    : transmittance_texel_split<float, P>(b, ix, iy, w, h, part);
    if (part == 0) out[t] = make_float4(v.x, v.y, v.z, 1.0f);
}

// Multiple scattering with each direction's march split over kSetMsParts waves: wave p of a
// texel's workgroup integrates part p of all 64 directions, waves 1.. leave their partial sums
// in LDS, and wave 0 combines them in order and reduces over the directions as the per-pass
// kernel does.
constexpr int kSetMsParts = 4;
__global__ __launch_bounds__(64 * kSetMsParts) void set_multiscattering_kernel(SkyhipAtmosphereBufferData b,
                                                                               const float4* __restrict__ T, int tw,
                                                                               int th, float4* __restrict__ out,
                                                                               int w, int h) {
    constexpr int P = kSetMsParts;
    __shared__ float part_sums[P - 1][9][64];
    const AtmDev A = make_atm(b);
    const int texel = blockIdx.x, part = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const MsRay ray = ms_ray(A, texel % w, texel / w, w, h, lane);
    int n = 0;
    for (float i = 0.5f; i < A.ms_steps; i += 1.0f) ++n;
    const int c = (n + P - 1) / P, first = min(part * c, n);
    const ScatterRay sc = scatter_ray<true>(A, ray.s, ray.dist, A.ms_steps);
    F3 L, Tr, Lf;
    scattered_luminance_part_ms(A, T, tw, th, ray.s, sc, 0.5f + float(first), max(0, min(c, n - first)), L, Tr, Lf);
    if (part > 0) {
        const float v[9] = {L.x, L.y, L.z, Tr.x, Tr.y, Tr.z, Lf.x, Lf.y, Lf.z};
        for (int k = 0; k < 9; ++k) part_sums[part - 1][k][lane] = v[k];
    }
    __syncthreads();
    if (part != 0) return;
    F3 aL = f3(0.0f), aT = f3(1.0f), aF = f3(0.0f);  // the parts after this one, combined
    for (int p = P - 1; p >= 1; --p) {
        const F3 pL = {part_sums[p - 1][0][lane], part_sums[p - 1][1][lane], part_sums[p - 1][2][lane]};
        const F3 pT = {part_sums[p - 1][3][lane], part_sums[p - 1][4][lane], part_sums[p - 1][5][lane]};
        const F3 pF = {part_sums[p - 1][6][lane], part_sums[p - 1][7][lane], part_sums[p - 1][8][lane]};
        if (p == P - 1) {
            aL = pL, aT = pT, aF = pF;
        } else {
            aL = pL + pT * aL;
            aF = pF + pT * aF;
            aT = pT * aT;
        }
    }
    L = L + Tr * aL;
    Lf = Lf + Tr * aF;
    Tr = Tr * aT;
    const float4 v = multiscattering_reduce(A, T, tw, th, ray, L, Tr, Lf);
    if (lane == 0) out[texel] = v;
}

// Blocks [0, sv_blocks) build the sky view, kSetSkyViewParts lanes per texel; the rest the
// aerial perspective, kSetAerialParts lanes per froxel; a wave holds 64 / parts whole groups (a
// lane left over idles) (no blue-noise dither, as the set's
// per-pass calls).
__global__ __launch_bounds__(256) void set_sky_aerial_kernel(
    SkyhipAtmosphereBufferData b, SkyhipAtmosphereRenderBufferData rbd, const float4* __restrict__ T, int tw, int th,
    const float4* __restrict__ MS, int mw, int mh, float4* __restrict__ sv_l, float4* __restrict__ sv_t, int sw,
    int sh, int sv_blocks, float4* __restrict__ ap_l, float4* __restrict__ ap_t, int aw, int ah, int ad) {
    const AtmDev A = make_atm(b);
    const RenderDev R = make_render(rbd);
    F3 L, Tr;
    if (int(blockIdx.x) < sv_blocks) {
        constexpr int P = kSetSkyViewParts, kPerWave = 64 / P;  // groups of P lanes within a wave
        const int g = blockIdx.x * 256 + threadIdx.x, lane = g % 64, slot = lane / P;
        const int part = lane % P;
        const int texel = (g / 64) * kPerWave + slot;
        const bool valid = slot < kPerWave && texel < sw * sh;
        const int t = valid ? texel : 0;  // every lane of a wave reaches the shuffles
        const SkyViewRay ray = sky_view_ray(A, R, nullptr, 1, t % sw, t / sw, sw, sh);
        scattered_luminance_split<P>(A, T, tw, th, MS, mw, mh, ray.s, ray.dist, R.sv_steps, ray.start_i,
                                     ray.dist > 0.0f, part, L, Tr);
        if (part == 0 && valid) {
            sv_l[t] = make_float4(L.x, L.y, L.z, 0.0f);
            sv_t[t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
        }
        return;
    }
    constexpr int P = kSetAerialParts, kPerWave = 64 / P;
    const long long g = (long long)(blockIdx.x - sv_blocks) * 256 + threadIdx.x;
    const int n_texels = aw * ah * ad;
    const int lane = int(g % 64), slot = lane / P, part = lane % P;
    const long long texel = (g / 64) * kPerWave + slot;
    const bool valid = slot < kPerWave && texel < n_texels;
    const int t = valid ? int(texel) : 0;
    const AerialRay ray = aerial_ray(A, R, rbd.inv_view_projection, nullptr, 1, t % aw, (t / aw) % ah, t / (aw * ah),
                                     aw, ah, ad);
    scattered_luminance_split<P>(A, T, tw, th, MS, mw, mh, ray.s, ray.marching_distance, R.ap_steps, ray.start_i,
                                 ray.marching_distance > 0.0f, part, L, Tr);
    if (part == 0 && valid) {
        ap_l[t] = make_float4(L.x, L.y, L.z, 0.0f);
        ap_t[t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Validation and launch helpers
// ---------------------------------------------------------------------------
constexpr int kMaxLutDim = 8192;

int validate_atmosphere(const SkyhipAtmosphereBufferData* b) {
    SKYHIP_CHECK(require(b != nullptr, SKYHIP_ERR_INVALID_ARGUMENT, "atmosphere block is NULL"));
    SKYHIP_CHECK(require(finite_block(reinterpret_cast<const float*>(b), sizeof(*b) / 4), SKYHIP_ERR_NON_FINITE,
                         "atmosphere block has a non-finite field"));
    SKYHIP_CHECK(require(b->bottom_radius > 0.0f && b->top_radius > b->bottom_radius, SKYHIP_ERR_NON_FINITE,
                         "atmosphere needs 0 < bottom_radius < top_radius"));
    SKYHIP_CHECK(require(b->transmittance_steps > 0.0f && b->transmittance_steps <= 4096.0f &&
                             b->multiscattering_steps > 0.0f && b->multiscattering_steps <= 4096.0f,
                         SKYHIP_ERR_INVALID_SIZE, "step counts must be in (0, 4096]"));
    return SKYHIP_OK;
}

int validate_render(const SkyhipAtmosphereRenderBufferData* r) {
    SKYHIP_CHECK(require(r != nullptr, SKYHIP_ERR_INVALID_ARGUMENT, "render block is NULL"));
    SKYHIP_CHECK(require(finite_block(reinterpret_cast<const float*>(r), sizeof(*r) / 4), SKYHIP_ERR_NON_FINITE,
                         "render block has a non-finite field"));
    SKYHIP_CHECK(require(r->sky_view_lut_steps >= 0.0f && r->sky_view_lut_steps <= 4096.0f &&
                             r->aerial_perspective_lut_steps >= 0.0f && r->aerial_perspective_lut_steps <= 4096.0f,
                         SKYHIP_ERR_INVALID_SIZE, "step counts must be in [0, 4096]"));
    return SKYHIP_OK;
}

int validate_lut(const void* p, int w, int h, const char* what) {
    if (p == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "%s pointer is NULL", what);
    if (w < 2 || h < 2 || w > kMaxLutDim || h > kMaxLutDim)
        return fail(SKYHIP_ERR_INVALID_SIZE, "%s size %dx%d outside [2, %d]", what, w, h, kMaxLutDim);
    return SKYHIP_OK;
}

int validate_blue_noise(const uint16_t* bn, int size) {
    if (bn == nullptr) return SKYHIP_OK;
    if (size < 1 || size > 4096 || (size & (size - 1)) != 0)
        return fail(SKYHIP_ERR_INVALID_SIZE, "blue-noise tile size %d is not a power of two in [1, 4096]", size);
    return SKYHIP_OK;
}

int launch_transmittance(hipStream_t st, const SkyhipAtmosphereBufferData& b, float* out, int w, int h) {
    hipLaunchKernelGGL(transmittance_kernel, dim3(div_up(w * h, 256)), dim3(256), 0, st, b,
                       reinterpret_cast<float4*>(out), w, h);
    SKYHIP_LAUNCH_CHECK("transmittance_kernel");
    return SKYHIP_OK;
}

int launch_multiscattering(hipStream_t st, const SkyhipAtmosphereBufferData& b, const float* T, int tw, int th,
                           float* out, int w, int h) {
    hipLaunchKernelGGL(multiscattering_kernel, dim3(div_up(w * h, 4)), dim3(256), 0, st, b,
                       reinterpret_cast<const float4*>(T), tw, th, reinterpret_cast<float4*>(out), w, h);
    SKYHIP_LAUNCH_CHECK("multiscattering_kernel");
    return SKYHIP_OK;
}

int launch_sky_view(hipStream_t st, const SkyhipAtmosphereBufferData& b, const SkyhipAtmosphereRenderBufferData& r,
                    const float* T, int tw, int th, const float* MS, int mw, int mh, const uint16_t* bn, int bn_size,
                    float* L, float* Tr, int w, int h) {
    hipLaunchKernelGGL(sky_view_kernel, dim3(div_up(w * h, 256)), dim3(256), 0, st, b, r,
                       reinterpret_cast<const float4*>(T), tw, th, reinterpret_cast<const float4*>(MS), mw, mh, bn,
                       bn_size, reinterpret_cast<float4*>(L), reinterpret_cast<float4*>(Tr), w, h);
    SKYHIP_LAUNCH_CHECK("sky_view_kernel");
    return SKYHIP_OK;
}

int launch_aerial_perspective(hipStream_t st, const SkyhipAtmosphereBufferData& b,
                              const SkyhipAtmosphereRenderBufferData& r, const float* T, int tw, int th,
                              const float* MS, int mw, int mh, const uint16_t* bn, int bn_size, float* L, float* Tr,
                              int w, int h, int d) {
    hipLaunchKernelGGL(aerial_perspective_kernel, dim3(div_up((long long)w * h * d * kApLanes, 256)), dim3(256), 0, st,
                       b, r,
                       reinterpret_cast<const float4*>(T), tw, th, reinterpret_cast<const float4*>(MS), mw, mh, bn,
                       bn_size, reinterpret_cast<float4*>(L), reinterpret_cast<float4*>(Tr), w, h, d);
    SKYHIP_LAUNCH_CHECK("aerial_perspective_kernel");
    return SKYHIP_OK;
}

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_transmittance_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                        float* transmittance_rgba, int32_t width, int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, width, height, "transmittance"));
    return launch_transmittance(ctx->stream, *atmosphere, transmittance_rgba, width, height);
}

extern "C" int skyhip_multiscattering_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                          const float* transmittance_rgba, int32_t transmittance_width,
                                          int32_t transmittance_height, float* multiscattering_rgba, int32_t width,
                                          int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, width, height, "multiscattering"));
    return launch_multiscattering(ctx->stream, *atmosphere, transmittance_rgba, transmittance_width,
                                  transmittance_height, multiscattering_rgba, width, height);
}

extern "C" int skyhip_atmosphere_update_luts(SkyhipContext* ctx, const SkyhipAtmosphereParameters* parameters,
                                             float* transmittance_rgba, int32_t transmittance_width,
                                             int32_t transmittance_height, float* multiscattering_rgba,
                                             int32_t multiscattering_width, int32_t multiscattering_height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SkyhipAtmosphereBufferData b;
    SKYHIP_CHECK(skyhip_atmosphere_assign_buffer_data(parameters, &b));
    SKYHIP_CHECK(validate_atmosphere(&b));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, multiscattering_width, multiscattering_height, "multiscattering"));
    SKYHIP_CHECK(launch_transmittance(ctx->stream, b, transmittance_rgba, transmittance_width, transmittance_height));
    return launch_multiscattering(ctx->stream, b, transmittance_rgba, transmittance_width, transmittance_height,
                                  multiscattering_rgba, multiscattering_width, multiscattering_height);
}

extern "C" int skyhip_sky_view_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                   const SkyhipAtmosphereRenderBufferData* render, const float* transmittance_rgba,
                                   int32_t transmittance_width, int32_t transmittance_height,
                                   const float* multiscattering_rgba, int32_t multiscattering_width,
                                   int32_t multiscattering_height, const uint16_t* blue_noise, int32_t blue_noise_size,
                                   float* luminance_rgba, float* sky_transmittance_rgba, int32_t width,
                                   int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, multiscattering_width, multiscattering_height, "multiscattering"));
    SKYHIP_CHECK(validate_blue_noise(blue_noise, blue_noise_size));
    SKYHIP_CHECK(validate_lut(luminance_rgba, width, height, "sky-view luminance"));
    SKYHIP_CHECK(validate_lut(sky_transmittance_rgba, width, height, "sky-view transmittance"));
    return launch_sky_view(ctx->stream, *atmosphere, *render, transmittance_rgba, transmittance_width,
                           transmittance_height, multiscattering_rgba, multiscattering_width, multiscattering_height,
                           blue_noise, blue_noise_size, luminance_rgba, sky_transmittance_rgba, width, height);
}

extern "C" int skyhip_aerial_perspective_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                             const SkyhipAtmosphereRenderBufferData* render,
                                             const float* transmittance_rgba, int32_t transmittance_width,
                                             int32_t transmittance_height, const float* multiscattering_rgba,
                                             int32_t multiscattering_width, int32_t multiscattering_height,
                                             const uint16_t* blue_noise, int32_t blue_noise_size,
                                             float* luminance_rgba, float* ap_transmittance_rgba, int32_t width,
                                             int32_t height, int32_t depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, multiscattering_width, multiscattering_height, "multiscattering"));
    SKYHIP_CHECK(validate_blue_noise(blue_noise, blue_noise_size));
    SKYHIP_CHECK(validate_lut(luminance_rgba, width, height, "aerial-perspective luminance"));
    SKYHIP_CHECK(validate_lut(ap_transmittance_rgba, width, height, "aerial-perspective transmittance"));
    if (depth < 2 || depth > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "aerial-perspective depth %d outside [2, 1024]", depth);
    return launch_aerial_perspective(ctx->stream, *atmosphere, *render, transmittance_rgba, transmittance_width,
                                     transmittance_height, multiscattering_rgba, multiscattering_width,
                                     multiscattering_height, blue_noise, blue_noise_size, luminance_rgba,
                                     ap_transmittance_rgba, width, height, depth);
}

extern "C" int skyhip_atmosphere_lut_set(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                         const SkyhipAtmosphereRenderBufferData* render, float* transmittance_rgba,
                                         int32_t transmittance_width, int32_t transmittance_height,
                                         float* multiscattering_rgba, int32_t multiscattering_width,
                                         int32_t multiscattering_height, float* sky_view_luminance_rgba,
                                         float* sky_view_transmittance_rgba, int32_t sky_view_width,
                                         int32_t sky_view_height, float* ap_luminance_rgba,
                                         float* ap_transmittance_rgba, int32_t ap_width, int32_t ap_height,
                                         int32_t ap_depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    const int tw = transmittance_width, th = transmittance_height, mw = multiscattering_width,
              mh = multiscattering_height;
    SKYHIP_CHECK(validate_lut(transmittance_rgba, tw, th, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, mw, mh, "multiscattering"));
    SKYHIP_CHECK(validate_lut(sky_view_luminance_rgba, sky_view_width, sky_view_height, "sky-view luminance"));
    SKYHIP_CHECK(validate_lut(sky_view_transmittance_rgba, sky_view_width, sky_view_height, "sky-view transmittance"));
    SKYHIP_CHECK(validate_lut(ap_luminance_rgba, ap_width, ap_height, "aerial-perspective luminance"));
    SKYHIP_CHECK(validate_lut(ap_transmittance_rgba, ap_width, ap_height, "aerial-perspective transmittance"));
    if (ap_depth < 2 || ap_depth > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "aerial-perspective depth outside [2, 1024]");
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(set_transmittance_kernel, dim3(div_up((long long)tw * th * kSetTransmittanceParts, 256)),
                       dim3(256), 0, st, *atmosphere, reinterpret_cast<float4*>(transmittance_rgba), tw, th);
    SKYHIP_LAUNCH_CHECK("set_transmittance_kernel");
    hipLaunchKernelGGL(set_multiscattering_kernel, dim3(mw * mh), dim3(64 * kSetMsParts), 0, st, *atmosphere,
                       reinterpret_cast<const float4*>(transmittance_rgba), tw, th,
                       reinterpret_cast<float4*>(multiscattering_rgba), mw, mh);
    SKYHIP_LAUNCH_CHECK("set_multiscattering_kernel");
    // Whole waves of 64 / parts groups each (a group never straddles two waves).
    const int sv_blocks = int(div_up(div_up((long long)sky_view_width * sky_view_height, 64 / kSetSkyViewParts) * 64, 256));
    const int ap_blocks =
        int(div_up(div_up((long long)ap_width * ap_height * ap_depth, 64 / kSetAerialParts) * 64, 256));
    hipLaunchKernelGGL(set_sky_aerial_kernel, dim3(sv_blocks + ap_blocks), dim3(256), 0, st, *atmosphere, *render,
                       reinterpret_cast<const float4*>(transmittance_rgba), tw, th,
                       reinterpret_cast<const float4*>(multiscattering_rgba), mw, mh,
                       reinterpret_cast<float4*>(sky_view_luminance_rgba),
                       reinterpret_cast<float4*>(sky_view_transmittance_rgba), sky_view_width, sky_view_height,
                       sv_blocks, reinterpret_cast<float4*>(ap_luminance_rgba),
                       reinterpret_cast<float4*>(ap_transmittance_rgba), ap_width, ap_height, ap_depth);
    SKYHIP_LAUNCH_CHECK("set_sky_aerial_kernel");
    return SKYHIP_OK;
}
