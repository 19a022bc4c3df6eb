// The volumetric cloud's shadow passes (VolumetricCloud::RenderShadow, VolumetricCloud.cpp:282-325):
//  - shadow_map_kernel: per texel of the 512^2 light-space map, the optical depth of the
//    cloud shell along -sun with 6-12 blue-noise-jittered steps through the material's
//    SampleSigmaT, stored as (mid-shell depth, transmittance) and blended 1:4 into the
//    previous frame's map reprojected (VolumetricCloudShadowMap.comp);
//  - shadow_blur_kernel: the separable 9-tap Gaussian, x then y, clamped at the edges
//    (VolumetricCloudShadowMapBlur.comp), both passes in one launch;
//  - shadow_froxel_kernel: per froxel column (viewport / 12), the running mean of the
//    blurred map's transmittance along the view ray over 128 slices, stored as R16 UNORM
//    (VolumetricCloudShadowFroxel.comp with SampleCloudShadowTransmittance,
//    VolumetricCloudShadowInterface.glsl:4-8); the slices' terms are evaluated in parallel and
//    summed per column in slice order.
// The arithmetic is FP32 as the shaders have it, except the shadow map's per-texel geometry
// (FP64, see shadow_map_kernel).
#include "cloud_device.h"

namespace skyhip {

int build_material(const SkyhipCloudMaterial* m, MaterialDev& d);
int validate_blue_noise(const uint16_t* bn, int size);

namespace {

// ProjectiveMul (shaders/Base/Common.glsl:7-10) in FP32, column-major m.
__device__ inline F3 projective_mul(const float* m, F3 v) {
    const float x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12];
    const float y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13];
    const float z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14];
    const float w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15];
    return {x / w, y / w, z / w};
}

// LINEAR lookup of an RG32F map at uv with CLAMP_TO_BORDER, border (1e10, 1) (the shadow-map
// sampler, VolumetricCloud.cpp:106-112).
__device__ inline float2 shadow_map_bilinear(const float2* __restrict__ map, int w, int h, float u, float v) {
    // Far outside the map (or NaN) every tap is border: clamping first keeps the integer
    // conversion in range without changing the result.
    const float fx = fminf(fmaxf(u * float(w) - 0.5f, -2.0f), float(w) + 1.0f);
    const float fy = fminf(fmaxf(v * float(h) - 0.5f, -2.0f), float(h) + 1.0f);
    const float x0f = floorf(fx), y0f = floorf(fy);
    const int x0 = int(x0f), y0 = int(y0f);
    const float ax = fx - x0f, ay = fy - y0f;
    float2 c[4];
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + (k & 1), y = y0 + (k >> 1);
        c[k] = (x < 0 || y < 0 || x >= w || y >= h) ? make_float2(1e10f, 1.0f) : map[y * w + x];
    }
    const float bx = 1.0f - ax, by = 1.0f - ay;
    return make_float2((c[0].x * bx + c[1].x * ax) * by + (c[2].x * bx + c[3].x * ax) * ay,
                       (c[0].y * bx + c[1].y * ax) * by + (c[2].y * bx + c[3].y * ax) * ay);
}

}  // namespace

struct ShadowMapArgs {
    MaterialDev mat;
    Shell S;
    float inv_light_vp[16], reproject[16];
    float frame_id;
    const uint16_t* blue_noise;
    int bn_size;
    const float2* previous;  // the previous frame's raw map, or null
    float2* out;
    int w, h;
};

// ProjectiveMul in FP64 of an FP32 matrix.
__device__ inline void projective_mul_d(const float* m, double x, double y, double z, double out[3]) {
    const double v[4] = {x, y, z, 1.0};
    double o[4];
    for (int row = 0; row < 4; ++row)
        o[row] = double(m[row]) * v[0] + double(m[4 + row]) * v[1] + double(m[8 + row]) * v[2] + double(m[12 + row]) * v[3];
    out[0] = o[0] / o[3];
    out[1] = o[1] / o[3];
    out[2] = o[2] / o[3];
}

// The per-texel geometry (the texel's origin on the light plane, its line's intersections
// with the cloud shell, the sample positions and the reprojection) runs in FP64: in FP32 the
// shader's r mu +- sqrt(discriminant) cancels between terms of ~2000 km and the +-50 km
// light-space matrix loses metres, so 9 % of the texels of the workload's map move by more
// than 1e-4 relative (the FP32 and FP64 oracles disagree there).  Densities are sampled in
// FP32 at the FP64 positions.
template <int kMat>
__global__ __launch_bounds__(256) void shadow_map_kernel(ShadowMapArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.w * a.h) return;
    const int x = i % a.w, y = i / a.w;
    const double u = (double(x) + 0.5) / double(a.w), v = (double(y) + 0.5) / double(a.h);
    double o[3];
    projective_mul_d(a.inv_light_vp, u * 2.0 - 1.0, v * 2.0 - 1.0, 0.0, o);
    const double dir[3] = {-double(a.S.sun.x), -double(a.S.sun.y), -double(a.S.sun.z)};
    const double R = a.S.R, cx = o[0], cy = o[1], cz = o[2] + R;
    const double r = sqrt(cx * cx + cy * cy + cz * cz);
    const double mu = (cx * dir[0] + cy * dir[1] + cz * dir[2]) / r;
    // LineShellFirstIntersect (:19-35)
    double t1 = 0.0, t2 = 0.0;
    const double rb = R + double(a.S.hb), rt = R + double(a.S.ht);
    const double disc_top = r * r * (mu * mu - 1.0) + rt * rt;
    if (disc_top > 0.0) {
        const double disc_bottom = r * r * (mu * mu - 1.0) + rb * rb;
        t1 = -r * mu - sqrt(disc_top);
        t2 = -r * mu + (disc_bottom >= 0.0 ? -sqrt(disc_bottom) : sqrt(disc_top));
    }
    double dist = fmax(t2 - t1, 0.0);
    dist = fmin(dist, 1.0 / cos(85.0 * 0.017453292519943295769) * double(a.S.ht - a.S.hb));
    double optical_depth = 0.0;
    if (dist > 0.0) {
        // mix(12, 6, |dir.z|) and its uint() in the shader's FP32
        const float az = fabsf(a.S.sun.z);
        const float steps = 12.0f * (1.0f - az) + 6.0f * az;
        const double step_size = dist / double(steps);
        const double noise = double(a.blue_noise[(y & 63) * a.bn_size + (x & 63)]) / 65535.0;
        const double jitter = noise + double(a.frame_id) * 0.61803398875;
        double t = t1 + step_size * (jitter - floor(jitter));
        for (uint32_t cnt = uint32_t(steps); cnt != 0u; --cnt, t += step_size) {
            const F3 pos = {float(o[0] + dir[0] * t), float(o[1] + dir[1] * t), float(o[2] + dir[2] * t)};
            optical_depth += double(sample_sigma_t<kMat>(a.mat, pos, height01(a.S, pos), a.S.cam)) * step_size;
        }
    }
    float2 res = make_float2(float(t1 * 0.5 + t2 * 0.5), float(exp(-optical_depth)));
    if (a.previous) {  // temporal blend with the reprojected previous map (:67-72)
        double p[3];
        projective_mul_d(a.reproject, u * 2.0 - 1.0, v * 2.0 - 1.0, 0.0, p);
        const double pu = p[0] * 0.5 + 0.5, pv = p[1] * 0.5 + 0.5;
        if (pu >= 0.5 / a.w && pu <= (a.w - 0.5) / a.w && pv >= 0.5 / a.h && pv <= (a.h - 0.5) / a.h) {
            const float2 pre = shadow_map_bilinear(a.previous, a.w, a.h, float(pu), float(pv));
            res = make_float2(pre.x * 0.8f + res.x * 0.2f, pre.y * 0.8f + res.y * 0.2f);
        }
    }
    a.out[i] = res;
}

// Both passes of the separable 9-tap Gaussian (weights of VolumetricCloudShadowMapBlur.comp:10),
// x then y, clamped at the edges, in one launch.  A 16x16 tile's workgroup computes the x pass
// for its rows and the 4 rows above and below them (rows past the map's edge clamped to it, as
// the y pass's taps are), keeps those in LDS, stores its own rows of them as the x-pass map
// `mid`, and runs the y pass from LDS.  Each value is the two-launch form's sum, the same
// multiply-adds in the same order.
constexpr int kBlurTile = 16, kBlurHalo = 4;
__global__ __launch_bounds__(256) void shadow_blur_kernel(const float2* __restrict__ in, float2* __restrict__ mid,
                                                          float2* __restrict__ out, int w, int h) {
    __shared__ float2 s_x[kBlurTile + 2 * kBlurHalo][kBlurTile];
    const float wt[5] = {0.227027f, 0.1945946f, 0.1216216f, 0.054054f, 0.016216f};
    const int x0 = int(blockIdx.x) * kBlurTile, y0 = int(blockIdx.y) * kBlurTile;
    for (int i = threadIdx.x; i < (kBlurTile + 2 * kBlurHalo) * kBlurTile; i += 256) {
        const int ly = i / kBlurTile, lx = i % kBlurTile;
        const int x = x0 + lx, yr = y0 - kBlurHalo + ly, y = min(max(yr, 0), h - 1);
        if (x >= w) continue;
        const float2 c = in[y * w + x];
        float rx = c.x * wt[0], ry = c.y * wt[0];
        for (int k = 1; k < 5; ++k) {
            const float2 a = in[y * w + max(x - k, 0)], b = in[y * w + min(x + k, w - 1)];
            rx += a.x * wt[k];
            ry += a.y * wt[k];
            rx += b.x * wt[k];
            ry += b.y * wt[k];
        }
        s_x[ly][lx] = make_float2(rx, ry);
        if (ly >= kBlurHalo && ly < kBlurHalo + kBlurTile && yr < h) mid[yr * w + x] = make_float2(rx, ry);
    }
    __syncthreads();
    const int lx = int(threadIdx.x) % kBlurTile, ly = int(threadIdx.x) / kBlurTile;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    const float2 c = s_x[ly + kBlurHalo][lx];
    float rx = c.x * wt[0], ry = c.y * wt[0];
    for (int k = 1; k < 5; ++k) {
        const float2 a = s_x[ly + kBlurHalo - k][lx], b = s_x[ly + kBlurHalo + k][lx];
        rx += a.x * wt[k];
        ry += a.y * wt[k];
        rx += b.x * wt[k];
        ry += b.y * wt[k];
    }
    out[y * w + x] = make_float2(rx, ry);
}

struct FroxelArgs {
    float inv_mvp[16], light_vp[16];
    F3 cam;
    float max_distance;
    const float2* map;
    int mw, mh;
    uint16_t* out;
    int w, h, d;
};

// A workgroup takes kFroxelCols consecutive columns of one froxel row.  The slices' terms are
// independent of each other, so its 256 threads evaluate them in parallel, kFroxelChunk slices
// at a time, into LDS; then one thread per column forms the running mean in slice order.  The
// slice distances come from a table that one thread fills with the shader's running
// t += step_size, and the column's thread leaves the running sums in the table for all threads to
// divide and store, so every term, sum and output is the one-thread-per-column loop's.  16
// columns a workgroup: the 160x90 froxel grid of a 1080p view gives 900 workgroups, 3.5 waves per
// SIMD (measured per shadow frame: 32 columns 0.0443 ms, 16 0.0405-0.0408, 8 0.0417).
constexpr int kFroxelCols = 16;
constexpr int kFroxelChunk = 128;
__global__ __launch_bounds__(256) void shadow_froxel_kernel(FroxelArgs a) {
    __shared__ float s_t[kFroxelChunk];
    __shared__ float s_term[kFroxelChunk][kFroxelCols];
    const int tiles_x = (a.w + kFroxelCols - 1) / kFroxelCols;
    const int y = int(blockIdx.x) / tiles_x, x = (int(blockIdx.x) % tiles_x) * kFroxelCols + int(threadIdx.x) % kFroxelCols;
    const int c = int(threadIdx.x) % kFroxelCols, g = int(threadIdx.x) / kFroxelCols;  // g in [0, 256 / kFroxelCols)
    const bool column = x < a.w;  // y < a.h by the grid
    const float u = (float(x) + 0.5f) / float(a.w), v = (float(y) + 0.5f) / float(a.h);
    const F3 frag = projective_mul(a.inv_mvp, F3{u * 2.0f - 1.0f, v * 2.0f - 1.0f, 0.0f});
    const float step_size = a.max_distance / float(a.d);
    F3 dir = frag - a.cam;
    dir = dir * (1.0f / sqrtf(dot(dir, dir)));
    float t = 0.5f * step_size, sum = 0.0f;  // t: thread 0's running distance; sum: the column threads'
    for (int z0 = 0; z0 < a.d; z0 += kFroxelChunk) {
        const int n = min(kFroxelChunk, a.d - z0);
        if (threadIdx.x == 0)
            for (int k = 0; k < n; ++k, t += step_size) s_t[k] = t;
        __syncthreads();
        if (column)
            for (int k = g; k < n; k += 256 / kFroxelCols) {
                const F3 pos = a.cam + dir * s_t[k];
                const F3 ndc = projective_mul(a.light_vp, pos);
                // SampleCloudShadowTransmittance (VolumetricCloudShadowInterface.glsl:4-8)
                const float2 dt = shadow_map_bilinear(a.map, a.mw, a.mh, ndc.x * 0.5f + 0.5f, ndc.y * 0.5f + 0.5f);
                const float s = clamp01((dt.x - ndc.z) * 2.0f);
                s_term[k][c] = dt.y * (1.0f - s) + 1.0f * s;
            }
        __syncthreads();
        if (column && g == 0)  // the running sum in slice order, left in place of the terms
            for (int k = 0; k < n; ++k) {
                sum += s_term[k][c];
                s_term[k][c] = sum;
            }
        __syncthreads();
        if (column)  // the means and their R16 UNORM stores, all threads
            for (int k = g; k < n; k += 256 / kFroxelCols) {
                const int z = z0 + k;
                const float vis = s_term[k][c] / float(z + 1);
                a.out[((long long)z * a.h + y) * a.w + x] = uint16_t(rintf(clamp01(vis) * 65535.0f));
            }
        __syncthreads();  // the next chunk overwrites the tables
    }
}

template <int kMat>
static void launch_shadow_map(hipStream_t st, const ShadowMapArgs& a) {
    hipLaunchKernelGGL(shadow_map_kernel<kMat>, dim3(div_up((long long)a.w * a.h, 256)), dim3(256), 0, st, a);
}

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_cloud_shadow_map(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                       const SkyhipCloudMaterial* material, const uint16_t* blue_noise,
                                       int32_t blue_noise_size, const float* previous_raw_rg32f,
                                       float* out_raw_rg32f, float* scratch_rg32f, float* out_blurred_rg32f,
                                       int32_t width, int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !out_raw_rg32f || !scratch_rg32f || !out_blurred_rg32f)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (width < 1 || height < 1 || width > 8192 || height > 8192)
        return fail(SKYHIP_ERR_INVALID_SIZE, "shadow map size outside [1, 8192]");
    if (!blue_noise) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "the shadow map needs the blue-noise tile");
    SKYHIP_CHECK(validate_blue_noise(blue_noise, blue_noise_size));
    if (blue_noise_size < 64) return fail(SKYHIP_ERR_INVALID_SIZE, "the shadow map reads a 64 x 64 blue-noise tile");
    if (!finite_block(common->uInvLightVP, 16) || !finite_block(common->uShadowMapReprojectMat, 16) ||
        !finite_block(common->uCameraPos, 3) || !finite_block(common->uSunDirection, 3))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud common block has a non-finite field");
    if (!(common->uTopAltitude > common->uBottomAltitude)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "empty shell");
    ShadowMapArgs a;
    std::memset(&a, 0, sizeof(a));
    SKYHIP_CHECK(build_material(material, a.mat));
    a.S.cam = f3(common->uCameraPos);
    a.S.sun = f3(common->uSunDirection);
    a.S.R = common->uEarthRadius;
    a.S.hb = common->uBottomAltitude;
    a.S.ht = common->uTopAltitude;
    a.S.inv_thickness = 1.0f / (a.S.ht - a.S.hb);
    std::memcpy(a.inv_light_vp, common->uInvLightVP, sizeof(a.inv_light_vp));
    std::memcpy(a.reproject, common->uShadowMapReprojectMat, sizeof(a.reproject));
    a.frame_id = common->uFrameID;
    a.blue_noise = blue_noise;
    a.bn_size = blue_noise_size;
    a.previous = reinterpret_cast<const float2*>(previous_raw_rg32f);
    a.out = reinterpret_cast<float2*>(out_raw_rg32f);
    a.w = width;
    a.h = height;
    switch (a.mat.kind) {
        case SKYHIP_MATERIAL_DEFAULT0: launch_shadow_map<SKYHIP_MATERIAL_DEFAULT0>(ctx->stream, a); break;
        case SKYHIP_MATERIAL_DEFAULT1: launch_shadow_map<SKYHIP_MATERIAL_DEFAULT1>(ctx->stream, a); break;
        case SKYHIP_MATERIAL_DEFAULT1_WORLEY:
            launch_shadow_map<SKYHIP_MATERIAL_DEFAULT1_WORLEY>(ctx->stream, a);
            break;
        case SKYHIP_MATERIAL_MINIMAL: launch_shadow_map<SKYHIP_MATERIAL_MINIMAL>(ctx->stream, a); break;
        case SKYHIP_MATERIAL_VOXEL: launch_shadow_map<SKYHIP_MATERIAL_VOXEL>(ctx->stream, a); break;
        default: launch_shadow_map<SKYHIP_MATERIAL_VOXEL_F16>(ctx->stream, a); break;
    }
    SKYHIP_LAUNCH_CHECK("shadow_map_kernel");
    hipLaunchKernelGGL(shadow_blur_kernel, dim3(div_up(width, kBlurTile), div_up(height, kBlurTile)), dim3(256), 0,
                       ctx->stream, reinterpret_cast<const float2*>(out_raw_rg32f), reinterpret_cast<float2*>(scratch_rg32f),
                       reinterpret_cast<float2*>(out_blurred_rg32f), width, height);
    SKYHIP_LAUNCH_CHECK("shadow_blur_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_shadow_froxel(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                          const float* shadow_map_rg32f, int32_t map_width, int32_t map_height,
                                          uint16_t* out_froxel_r16, int32_t width, int32_t height, int32_t depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !shadow_map_rg32f || !out_froxel_r16) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (map_width < 1 || map_height < 1 || width < 1 || height < 1 || depth < 1 || width > 8192 ||
        height > 8192 || depth > 1024)
        return fail(SKYHIP_ERR_INVALID_SIZE, "froxel or shadow-map size out of range");
    if (!finite_block(common->uInvMVP, 16) || !finite_block(common->uLightVP, 16) ||
        !finite_block(common->uCameraPos, 3) || !(common->uShadowFroxelMaxDistance > 0.0f))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud common block: non-finite matrices or froxel distance <= 0");
    FroxelArgs a;
    std::memcpy(a.inv_mvp, common->uInvMVP, sizeof(a.inv_mvp));
    std::memcpy(a.light_vp, common->uLightVP, sizeof(a.light_vp));
    a.cam = f3(common->uCameraPos);
    a.max_distance = common->uShadowFroxelMaxDistance;
    a.map = reinterpret_cast<const float2*>(shadow_map_rg32f);
    a.mw = map_width;
    a.mh = map_height;
    a.out = out_froxel_r16;
    a.w = width;
    a.h = height;
    a.d = depth;
    hipLaunchKernelGGL(shadow_froxel_kernel, dim3(unsigned(div_up(width, kFroxelCols) * height)), dim3(256), 0,
                       ctx->stream, a);
    SKYHIP_LAUNCH_CHECK("shadow_froxel_kernel");
    return SKYHIP_OK;
}
