// The volumetric cloud's quarter-resolution pipeline around the raymarch
// (VolumetricCloud::Render, VolumetricCloud.cpp:327-422):
//  - checkerboard_depth_kernel (W/2 x H/2): per 2x2 block of the depth buffer its max on
//    one checkerboard colour and its min on the other (CheckerboardGen.comp);
//  - shading_index_kernel (W/4 x H/4): which of the 4 checkerboard texels of a tile the
//    raymarch shades this frame, and its linear depth (VolumetricCloudIndexGen.comp);
//  - [skyhip_cloud_raymarch with the index: the quarter-resolution render]
//  - reconstruct_kernel (W/2 x H/2): the depth-nearest rendered neighbour's colour box
//    clamps the reprojected history, blended 1:4 with the new sample where it was shaded
//    (VolumetricCloudReconstruct.comp, Reinhard space);
//  - upscale_kernel (W x H): depth-aware 2x upscale composited over the HDR image
//    (VolumetricCloudUpscale.comp).
// All four passes are selections, comparisons and short FP32 expressions: they run with FMA
// contraction off, so that each result is the one an FP32 evaluation of the shader's
// expressions, in their order, gives.
#include "cloud_device.h"

namespace skyhip {

namespace {

constexpr float kMinTransmittance = 0.01f;  // VolumetricCloudCommon.glsl:30

__device__ inline float linear_depth(const float p[2], float depth) {  // DepthToLinearDepth (:32-34)
#pragma clang fp contract(off)
    return 1.0f / (p[0] - p[1] * depth);
}
__device__ inline int offset_x(uint32_t index) { return int(((index + 1u) >> 1) & 1u); }  // IndexToOffset (:42-52)
__device__ inline int offset_y(uint32_t index) { return int(((index + 2u) >> 1) & 1u); }

__device__ inline float4 load_rgba16f(const ushort4* p, int i) {
    const ushort4 v = p[i];
    return make_float4(float(__builtin_bit_cast(_Float16, v.x)), float(__builtin_bit_cast(_Float16, v.y)),
                       float(__builtin_bit_cast(_Float16, v.z)), float(__builtin_bit_cast(_Float16, v.w)));
}
__device__ inline ushort4 store_rgba16f(float4 c) {
    ushort4 o;
    o.x = __half_as_ushort(__float2half_rn(c.x));
    o.y = __half_as_ushort(__float2half_rn(c.y));
    o.z = __half_as_ushort(__float2half_rn(c.z));
    o.w = __half_as_ushort(__float2half_rn(c.w));
    return o;
}
__device__ inline float4 reinhard(float4 c) {
#pragma clang fp contract(off)
    return make_float4(c.x / (1.0f + c.x), c.y / (1.0f + c.y), c.z / (1.0f + c.z), c.w);
}
__device__ inline float4 inverse_reinhard(float4 c) {
#pragma clang fp contract(off)
    return make_float4(c.x / (1.0f - c.x), c.y / (1.0f - c.y), c.z / (1.0f - c.z), c.w);
}
__device__ inline int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// ProjectiveMul (Common.glsl:7-10), FP32 in the shader's order.
__device__ inline F3 projective_mul_f(const float* m, F3 v) {
#pragma clang fp contract(off)
    const float x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12];
    const float y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13];
    const float z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14];
    const float w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15];
    return {x / w, y / w, z / w};
}

}  // namespace

// CheckerboardGen.comp: out (W/2 x H/2) from depth (W x H); textureGather at the block centre.
__global__ __launch_bounds__(256) void checkerboard_depth_kernel(const float* __restrict__ depth, int w, int h,
                                                                 float* __restrict__ out) {
    const int hw = w / 2, hh = h / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= hw * hh) return;
    const int x = i % hw, y = i / hw;
    const float a = depth[(2 * y) * w + 2 * x], b = depth[(2 * y) * w + 2 * x + 1];
    const float c = depth[(2 * y + 1) * w + 2 * x], d = depth[(2 * y + 1) * w + 2 * x + 1];
    const bool bmax = (x & 1) == (y & 1);
    out[i] = bmax ? fmaxf(fmaxf(a, b), fmaxf(c, d)) : fminf(fminf(a, b), fminf(c, d));
}

struct IndexArgs {
    float lin[2];
    uint32_t base;
    const float* cb;  // checkerboard depth, W/2 x H/2
    int cw, ch;
    float2* out;  // W/4 x H/4: (index, linear depth)
};

// VolumetricCloudIndexGen.comp.  The neighbour tiles' fetches outside the checkerboard read 0,
// the value of an out-of-range texelFetch under robust access.
__global__ __launch_bounds__(256) void shading_index_kernel(IndexArgs a) {
#pragma clang fp contract(off)
    const int qw = a.cw / 2, qh = a.ch / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= qw * qh) return;
    const int px = i % qw, py = i / qw;
    // textureGather order: (0, 1), (1, 1), (1, 0), (0, 0) = IndexToOffset(0..3)
    float tile[4];
    for (int k = 0; k < 4; ++k) tile[k] = a.cb[(2 * py + offset_y(k)) * a.cw + 2 * px + offset_x(k)];
    uint32_t nearest = 0, farthest = 0;
    for (uint32_t k = 1; k < 4; ++k) {
        nearest = tile[k] < tile[nearest] ? k : nearest;
        farthest = tile[k] > tile[farthest] ? k : farthest;
    }
    const float near_lin = linear_depth(a.lin, tile[nearest]), far_lin = linear_depth(a.lin, tile[farthest]);
    uint32_t close_near = 0, close_far = 0;
    const int ox[8] = {0, 1, 1, 1, 0, -1, -1, -1}, oy[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    for (int k = 0; k < 8; ++k) {
        const int tx = px + ox[k], ty = py + oy[k];
        const uint32_t idx = (a.base + uint32_t((tx + ty) & 1)) & 3u;  // PosToIndex
        const int cx = tx * 2 + offset_x(idx), cy = ty * 2 + offset_y(idx);
        const float d = (cx < 0 || cy < 0 || cx >= a.cw || cy >= a.ch) ? 0.0f : a.cb[cy * a.cw + cx];
        const float lin = linear_depth(a.lin, d);
        const float max_delta = lin * 0.25f;
        if (fabsf(lin - near_lin) < max_delta) ++close_near;
        if (fabsf(lin - far_lin) < max_delta) ++close_far;
    }
    const uint32_t index = close_near == 0 ? nearest : close_far == 0 ? farthest : (a.base + uint32_t((px + py) & 1)) & 3u;
    a.out[i] = make_float2(float(index), linear_depth(a.lin, tile[index]));
}

struct ReconstructArgs {
    float lin[2];
    float inv_mvp[16], reproject[16];
    F3 cam;
    const float* cb;            // W/2 x H/2
    const float2* index;        // W/4 x H/4
    const ushort4* render;      // W/4 x H/4 RGBA16F
    const float* distance;      // W/4 x H/4
    const ushort4* previous;    // W/2 x H/2 RGBA16F, or null (first frame: the new samples count fully)
    ushort4* out;               // W/2 x H/2
    int hw, hh;
};

// Bilinear CLAMP_TO_EDGE lookup of an RGBA16F image (the linear no-mipmap sampler).
__device__ inline float4 rgba16f_bilinear(const ushort4* img, int w, int h, float u, float v) {
#pragma clang fp contract(off)
    const float fx = u * float(w) - 0.5f, fy = v * float(h) - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    const int x0 = int(fminf(fmaxf(x0f, -1.0f), float(w))), y0 = int(fminf(fmaxf(y0f, -1.0f), float(h)));
    const int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    const int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    const float4 c00 = load_rgba16f(img, ya * w + xa), c10 = load_rgba16f(img, ya * w + xb);
    const float4 c01 = load_rgba16f(img, yb * w + xa), c11 = load_rgba16f(img, yb * w + xb);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    return make_float4((c00.x * bx + c10.x * ax) * by + (c01.x * bx + c11.x * ax) * ay,
                       (c00.y * bx + c10.y * ax) * by + (c01.y * bx + c11.y * ax) * ay,
                       (c00.z * bx + c10.z * ax) * by + (c01.z * bx + c11.z * ax) * ay,
                       (c00.w * bx + c10.w * ax) * by + (c01.w * bx + c11.w * ax) * ay);
}

__device__ inline float4 min4(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), fminf(a.w, b.w));
}
__device__ inline float4 max4(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

// VolumetricCloudReconstruct.comp (USE_TONEMAPPING 1, USE_VARIANCE 0).
__global__ __launch_bounds__(256) void reconstruct_kernel(ReconstructArgs a) {
#pragma clang fp contract(off)
    const int qw = a.hw / 2, qh = a.hh / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.hw * a.hh) return;
    const int x = i % a.hw, y = i / a.hw, bx = x >> 1, by = y >> 1;
    const float u = (float(x) + 0.5f) / float(a.hw), v = (float(y) + 0.5f) / float(a.hh);
    const float depth = a.cb[i];
    const float lin = linear_depth(a.lin, depth);
    const int ox[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1}, oy[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
    float rendered_lin[9], delta[9];
    float min_delta = 1e10f;
    int nearest = 0;
    for (int k = 0; k < 9; ++k) {
        const int qx = clampi(bx + ox[k], 0, qw - 1), qy = clampi(by + oy[k], 0, qh - 1);
        rendered_lin[k] = a.index[qy * qw + qx].y;
        delta[k] = fabsf(rendered_lin[k] - lin);
        if (delta[k] < min_delta) {
            min_delta = delta[k];
            nearest = k;
        }
    }
    const float4 near_c = reinhard(load_rgba16f(
        a.render, clampi(by + oy[nearest], 0, qh - 1) * qw + clampi(bx + ox[nearest], 0, qw - 1)));
    float4 lo = near_c, hi = near_c;
    for (int k = 0; k < 9; ++k) {
        const float4 c = reinhard(load_rgba16f(a.render, clampi(by + oy[k], 0, qh - 1) * qw + clampi(bx + ox[k], 0, qw - 1)));
        if (delta[k] < rendered_lin[k] * 0.3f ||
            fabsf(c.w - near_c.w) / fmaxf(1e-6f, 1.0f - fmaxf(c.w, near_c.w)) < 0.2f) {
            lo = min4(lo, c);
            hi = max4(hi, c);
        }
    }
    const float4 rendered = reinhard(load_rgba16f(a.render, by * qw + bx));
    const F3 frag = projective_mul_f(a.inv_mvp, F3{u * 2.0f - 1.0f, v * 2.0f - 1.0f, depth * 2.0f - 1.0f});
    F3 dir = frag - a.cam;
    dir = dir * (1.0f / sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z));
    const float dist = a.distance[by * qw + bx];
    const F3 cloud_pos = a.cam + dir * dist;
    const F3 pre_ndc = projective_mul_f(a.reproject, cloud_pos);
    const float pu = pre_ndc.x * 0.5f + 0.5f, pv = pre_ndc.y * 0.5f + 0.5f;
    const bool out_of_screen = !a.previous || fmaxf(fabsf(pre_ndc.x), fabsf(pre_ndc.y)) > 1.0f;
    float4 pre = a.previous ? reinhard(rgba16f_bilinear(a.previous, a.hw, a.hh, pu, pv)) : rendered;
    pre = min4(max4(pre, lo), hi);  // clamp(pre, aabb_min, aabb_max)
    const uint32_t ridx = uint32_t(int(a.index[by * qw + bx].x));
    const bool is_rendered = (x & 1) == offset_x(ridx) && (y & 1) == offset_y(ridx);
    const float wgt = out_of_screen ? 1.0f : is_rendered ? 0.2f : 0.0f;
    const float iw = 1.0f - wgt;
    const float4 r = make_float4(pre.x * iw + rendered.x * wgt, pre.y * iw + rendered.y * wgt,
                                 pre.z * iw + rendered.z * wgt, pre.w * iw + rendered.w * wgt);  // mix
    a.out[i] = store_rgba16f(inverse_reinhard(r));
}

struct UpscaleArgs {
    float lin[2];
    const float* cb;             // W/2 x H/2
    const float* depth;          // W x H
    const ushort4* reconstruct;  // W/2 x H/2
    ushort4* hdr;                // W x H, composited in place
    int w, h;
};

// VolumetricCloudUpscale.comp.
__global__ __launch_bounds__(256) void upscale_kernel(UpscaleArgs a) {
#pragma clang fp contract(off)
    const int hw = a.w / 2, hh = a.h / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.w * a.h) return;
    const int x = i % a.w, y = i / a.w;
    const float lin = linear_depth(a.lin, a.depth[i]);
    const int ox[4] = {-1, 1, 1, -1}, oy[4] = {1, 1, -1, -1};  // textureGather order
    float4 n[4];
    float min_delta = 1e10f;
    int nearest = 0;
    bool edge = false;
    for (int k = 0; k < 4; ++k) {
        const int hx = (x + ox[k]) >> 1, hy = (y + oy[k]) >> 1;
        n[k] = load_rgba16f(a.reconstruct, clampi(hy, 0, hh - 1) * hw + clampi(hx, 0, hw - 1));
        const float nd = a.cb[clampi(hy, 0, hh - 1) * hw + clampi(hx, 0, hw - 1)];
        const float d = fabsf(lin - linear_depth(a.lin, nd));
        if (d < min_delta) {
            nearest = k;
            min_delta = d;
        }
        if (d > lin * 0.1f) edge = true;
    }
    const float max_a = fmaxf(fmaxf(n[0].w, n[1].w), fmaxf(n[2].w, n[3].w));
    const float min_a = fminf(fminf(n[0].w, n[1].w), fminf(n[2].w, n[3].w));
    float4 up;
    if (edge && (max_a - min_a) / fmaxf(1e-6f, 1.0f - min_a) > 0.2f) {
        up = n[nearest];
    } else {
        up = make_float4(((n[0].x + n[1].x) + (n[2].x + n[3].x)) * 0.25f, ((n[0].y + n[1].y) + (n[2].y + n[3].y)) * 0.25f,
                         ((n[0].z + n[1].z) + (n[2].z + n[3].z)) * 0.25f, ((n[0].w + n[1].w) + (n[2].w + n[3].w)) * 0.25f);
    }
    float4 color = load_rgba16f(a.hdr, i);
    const float t = up.w <= kMinTransmittance ? 0.0f : up.w;
    color.x = color.x * t + up.x;
    color.y = color.y * t + up.y;
    color.z = color.z * t + up.z;
    a.hdr[i] = store_rgba16f(color);
}

}  // namespace skyhip

using namespace skyhip;

namespace {
int check_viewport(int w, int h) {
    if (w < 4 || h < 4 || w > 16384 || h > 16384 || (w % 4) != 0 || (h % 4) != 0)
        return fail(SKYHIP_ERR_INVALID_SIZE, "viewport must be a multiple of 4 in [4, 16384]");
    return SKYHIP_OK;
}
}  // namespace

extern "C" int skyhip_cloud_checkerboard_depth(SkyhipContext* ctx, const float* depth, int32_t width, int32_t height,
                                               float* out_checkerboard) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!depth || !out_checkerboard) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(check_viewport(width, height));
    hipLaunchKernelGGL(checkerboard_depth_kernel, dim3(div_up((long long)(width / 2) * (height / 2), 256)), dim3(256),
                       0, ctx->stream, depth, width, height, out_checkerboard);
    SKYHIP_LAUNCH_CHECK("checkerboard_depth_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_shading_index(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                          const float* checkerboard, int32_t width, int32_t height,
                                          float* out_index_linear_depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !checkerboard || !out_index_linear_depth) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(check_viewport(width, height));
    if (!finite_block(common->uLinearDepthParam, 2)) return fail(SKYHIP_ERR_NON_FINITE, "uLinearDepthParam");
    IndexArgs a;
    a.lin[0] = common->uLinearDepthParam[0];
    a.lin[1] = common->uLinearDepthParam[1];
    a.base = common->uBaseShadingIndex;
    a.cb = checkerboard;
    a.cw = width / 2;
    a.ch = height / 2;
    a.out = reinterpret_cast<float2*>(out_index_linear_depth);
    hipLaunchKernelGGL(shading_index_kernel, dim3(div_up((long long)(width / 4) * (height / 4), 256)), dim3(256), 0,
                       ctx->stream, a);
    SKYHIP_LAUNCH_CHECK("shading_index_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_reconstruct(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                        const float* checkerboard, const float* index_linear_depth,
                                        const void* render_rgba16f, const float* cloud_distance,
                                        const void* previous_rgba16f, void* out_rgba16f, int32_t width,
                                        int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !checkerboard || !index_linear_depth || !render_rgba16f || !cloud_distance || !out_rgba16f)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(check_viewport(width, height));
    if (!finite_block(common->uInvMVP, 16) || !finite_block(common->uReprojectMat, 16) ||
        !finite_block(common->uCameraPos, 3) || !finite_block(common->uLinearDepthParam, 2))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud common block has a non-finite field");
    if (previous_rgba16f == out_rgba16f) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "previous and output alias");
    ReconstructArgs a;
    a.lin[0] = common->uLinearDepthParam[0];
    a.lin[1] = common->uLinearDepthParam[1];
    std::memcpy(a.inv_mvp, common->uInvMVP, sizeof(a.inv_mvp));
    std::memcpy(a.reproject, common->uReprojectMat, sizeof(a.reproject));
    a.cam = f3(common->uCameraPos);
    a.cb = checkerboard;
    a.index = reinterpret_cast<const float2*>(index_linear_depth);
    a.render = reinterpret_cast<const ushort4*>(render_rgba16f);
    a.distance = cloud_distance;
    a.previous = reinterpret_cast<const ushort4*>(previous_rgba16f);
    a.out = reinterpret_cast<ushort4*>(out_rgba16f);
    a.hw = width / 2;
    a.hh = height / 2;
    hipLaunchKernelGGL(reconstruct_kernel, dim3(div_up((long long)a.hw * a.hh, 256)), dim3(256), 0, ctx->stream, a);
    SKYHIP_LAUNCH_CHECK("reconstruct_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_upscale(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                    const float* checkerboard, const float* depth, const void* reconstruct_rgba16f,
                                    void* hdr_rgba16f, int32_t width, int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !checkerboard || !depth || !reconstruct_rgba16f || !hdr_rgba16f)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(check_viewport(width, height));
    if (!finite_block(common->uLinearDepthParam, 2)) return fail(SKYHIP_ERR_NON_FINITE, "uLinearDepthParam");
    UpscaleArgs a;
    a.lin[0] = common->uLinearDepthParam[0];
    a.lin[1] = common->uLinearDepthParam[1];
    a.cb = checkerboard;
    a.depth = depth;
    a.reconstruct = reinterpret_cast<const ushort4*>(reconstruct_rgba16f);
    a.hdr = reinterpret_cast<ushort4*>(hdr_rgba16f);
    a.w = width;
    a.h = height;
    hipLaunchKernelGGL(upscale_kernel, dim3(div_up((long long)width * height, 256)), dim3(256), 0, ctx->stream, a);
    SKYHIP_LAUNCH_CHECK("upscale_kernel");
    return SKYHIP_OK;
}
