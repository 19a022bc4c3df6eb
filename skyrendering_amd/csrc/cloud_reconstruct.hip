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

__device__ inline float4 unpack_rgba16f(ushort4 v) {
    return make_float4(float(__builtin_bit_cast(_Float16, v.x)), float(__builtin_bit_cast(_Float16, v.y)),
                       float(__builtin_bit_cast(_Float16, v.z)), float(__builtin_bit_cast(_Float16, v.w)));
}
__device__ inline float4 load_rgba16f(const ushort4* p, int i) { return unpack_rgba16f(p[i]); }
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

// VolumetricCloudIndexGen.comp's decision for tile (px, py) of a cw x ch checkerboard, from the
// tile's 4 texels in textureGather order ((0, 1), (1, 1), (1, 0), (0, 0) = IndexToOffset(0..3))
// and `fetch(cx, cy)`, a checkerboard texel inside the image.  The neighbour tiles' fetches
// outside the checkerboard read 0, the value of an out-of-range texelFetch under robust access.
template <class Fetch>
__device__ inline float2 shading_index(const float lin_p[2], uint32_t base, int px, int py, int cw, int ch,
                                       const float tile[4], const Fetch& fetch) {
#pragma clang fp contract(off)
    uint32_t nearest = 0, farthest = 0;
    float near_d = tile[0], far_d = tile[0];
#pragma unroll
    for (uint32_t k = 1; k < 4; ++k) {
        if (tile[k] < near_d) nearest = k, near_d = tile[k];
        if (tile[k] > far_d) farthest = k, far_d = tile[k];
    }
    const float near_lin = linear_depth(lin_p, near_d), far_lin = linear_depth(lin_p, far_d);
    uint32_t close_near = 0, close_far = 0;
    const int ox[8] = {0, 1, 1, 1, 0, -1, -1, -1}, oy[8] = {1, 1, 0, -1, -1, -1, 0, 1};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int tx = px + ox[k], ty = py + oy[k];
        const uint32_t idx = (base + uint32_t((tx + ty) & 1)) & 3u;  // PosToIndex
        const int cx = tx * 2 + offset_x(idx), cy = ty * 2 + offset_y(idx);
        const float d = (cx < 0 || cy < 0 || cx >= cw || cy >= ch) ? 0.0f : fetch(cx, cy);
        const float lin = linear_depth(lin_p, d);
        const float max_delta = lin * 0.25f;
        if (fabsf(lin - near_lin) < max_delta) ++close_near;
        if (fabsf(lin - far_lin) < max_delta) ++close_far;
    }
    const uint32_t index = close_near == 0 ? nearest : close_far == 0 ? farthest : (base + uint32_t((px + py) & 1)) & 3u;
    const float chosen = index == 0 ? tile[0] : index == 1 ? tile[1] : index == 2 ? tile[2] : tile[3];
    return make_float2(float(index), linear_depth(lin_p, chosen));
}

struct IndexArgs {
    float lin[2];
    uint32_t base;
    const float* cb;  // checkerboard depth, W/2 x H/2
    int cw, ch;
    float2* out;  // W/4 x H/4: (index, linear depth)
};

__global__ __launch_bounds__(256) void shading_index_kernel(IndexArgs a) {
    const int qw = a.cw / 2, qh = a.ch / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= qw * qh) return;
    const int px = i % qw, py = i / qw;
    float tile[4];
    for (int k = 0; k < 4; ++k) tile[k] = a.cb[(2 * py + offset_y(k)) * a.cw + 2 * px + offset_x(k)];
    const float* cb = a.cb;
    const int cw = a.cw;
    a.out[i] = shading_index(a.lin, a.base, px, py, a.cw, a.ch, tile,
                             [cb, cw](int cx, int cy) { return cb[cy * cw + cx]; });
}

struct CheckerboardIndexArgs {
    float lin[2];
    uint32_t base;
    const float* depth;  // W x H
    int w, h;
    float* cb;    // out: W/2 x H/2
    float2* out;  // out: W/4 x H/4
};

// Checkerboard texel (cx, cy) from the depth buffer: the max of its 2x2 block where cx and cy
// have the same parity, the min otherwise.
__device__ inline float checkerboard_texel(const float* __restrict__ depth, int w, int cx, int cy) {
    const float2 r0 = *reinterpret_cast<const float2*>(depth + (2 * cy) * w + 2 * cx);
    const float2 r1 = *reinterpret_cast<const float2*>(depth + (2 * cy + 1) * w + 2 * cx);
    return (cx & 1) == (cy & 1) ? fmaxf(fmaxf(r0.x, r0.y), fmaxf(r1.x, r1.y))
                                : fminf(fminf(r0.x, r0.y), fminf(r1.x, r1.y));
}

// CheckerboardGen.comp and VolumetricCloudIndexGen.comp in one pass, one thread per 2x2 tile of
// the checkerboard (a 4x4 block of the depth buffer, read as 4 float4 rows): the tile's 4
// texels are written to the checkerboard and the neighbour tiles' scheduled texels are formed
// from the depth buffer again (L2 hits), so no pass waits for another's output.  The values
// are the two passes' (max and min are exact).
__global__ __launch_bounds__(256) void checkerboard_index_kernel(CheckerboardIndexArgs a) {
    const int qw = a.w / 4, qh = a.h / 4, cw = a.w / 2, ch = a.h / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= qw * qh) return;
    const int px = i % qw, py = i / qw;
    const float* row = a.depth + (4 * py) * a.w + 4 * px;
    const float4 r0 = *reinterpret_cast<const float4*>(row);
    const float4 r1 = *reinterpret_cast<const float4*>(row + a.w);
    const float4 r2 = *reinterpret_cast<const float4*>(row + 2 * a.w);
    const float4 r3 = *reinterpret_cast<const float4*>(row + 3 * a.w);
    // Texel (2px + i, 2py + j): parity (i, j), so max where i == j.
    const float c00 = fmaxf(fmaxf(r0.x, r0.y), fmaxf(r1.x, r1.y));
    const float c10 = fminf(fminf(r0.z, r0.w), fminf(r1.z, r1.w));
    const float c01 = fminf(fminf(r2.x, r2.y), fminf(r3.x, r3.y));
    const float c11 = fmaxf(fmaxf(r2.z, r2.w), fmaxf(r3.z, r3.w));
    const float tile[4] = {c01, c11, c10, c00};  // IndexToOffset(0..3) = (0, 1), (1, 1), (1, 0), (0, 0)
    const float* depth = a.depth;
    const int w = a.w;
    // The neighbour reads come before any store, so all of the thread's loads issue together.
    const float2 index = shading_index(a.lin, a.base, px, py, cw, ch, tile,
                                       [depth, w](int cx, int cy) { return checkerboard_texel(depth, w, cx, cy); });
    *reinterpret_cast<float2*>(a.cb + (2 * py) * cw + 2 * px) = make_float2(c00, c10);
    *reinterpret_cast<float2*>(a.cb + (2 * py + 1) * cw + 2 * px) = make_float2(c01, c11);
    a.out[i] = index;
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

// The quarter-resolution tiles of the reconstruction read from global memory: (index, linear
// depth) and the Reinhard-mapped render sample at clamped tile coordinates.
struct GlobalTiles {
    const float2* index;
    const ushort4* render;
    int qw;
    __device__ float2 idx(int qx, int qy) const { return index[qy * qw + qx]; }
    __device__ float4 color(int qx, int qy) const { return reinhard(load_rgba16f(render, qy * qw + qx)); }
};

// VolumetricCloudReconstruct.comp (USE_TONEMAPPING 1, USE_VARIANCE 0) for half-resolution pixel
// (x, y), in two steps so that a thread can hold several pixels' memory reads in flight:
// reconstruct_prepare takes the pixel's checkerboard depth and cloud distance and forms the
// colour box, the new sample, the blend weight and where the history is read;
// reconstruct_finish reads the history and blends.
struct ReconstructStage {
    float4 lo, hi, rendered;
    float pu, pv, wgt;
};

template <class Tiles>
__device__ inline void reconstruct_prepare(const ReconstructArgs& a, const Tiles& t, int x, int y, float depth,
                                           float lin, float dist, ReconstructStage& st) {
#pragma clang fp contract(off)
    const int qw = a.hw / 2, qh = a.hh / 2;
    const int bx = x >> 1, by = y >> 1;
    const float u = (float(x) + 0.5f) / float(a.hw), v = (float(y) + 0.5f) / float(a.hh);
    const int ox[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1}, oy[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
    float rendered_lin[9], delta[9];
    float min_delta = 1e10f;
    int near_qx = bx, near_qy = by;  // kOffsets[nearest_i], clamped
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int qx = clampi(bx + ox[k], 0, qw - 1), qy = clampi(by + oy[k], 0, qh - 1);
        rendered_lin[k] = t.idx(qx, qy).y;
        delta[k] = fabsf(rendered_lin[k] - lin);
        if (delta[k] < min_delta) {
            min_delta = delta[k];
            near_qx = qx;
            near_qy = qy;
        }
    }
    const float4 near_c = t.color(near_qx, near_qy);
    float4 lo = near_c, hi = near_c;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float4 c = t.color(clampi(bx + ox[k], 0, qw - 1), clampi(by + oy[k], 0, qh - 1));
        // The shader's || evaluates the transmittance ratio only where the depth test fails.
        bool in_box = delta[k] < rendered_lin[k] * 0.3f;
        if (!in_box) in_box = fabsf(c.w - near_c.w) / fmaxf(1e-6f, 1.0f - fmaxf(c.w, near_c.w)) < 0.2f;
        if (in_box) {
            lo = min4(lo, c);
            hi = max4(hi, c);
        }
    }
    st.lo = lo;
    st.hi = hi;
    st.rendered = t.color(bx, by);
    const F3 frag = projective_mul_f(a.inv_mvp, F3{u * 2.0f - 1.0f, v * 2.0f - 1.0f, depth * 2.0f - 1.0f});
    F3 dir = frag - a.cam;
    dir = dir * (1.0f / sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z));
    const F3 cloud_pos = a.cam + dir * dist;
    const F3 pre_ndc = projective_mul_f(a.reproject, cloud_pos);
    st.pu = pre_ndc.x * 0.5f + 0.5f;
    st.pv = pre_ndc.y * 0.5f + 0.5f;
    const bool out_of_screen = !a.previous || fmaxf(fabsf(pre_ndc.x), fabsf(pre_ndc.y)) > 1.0f;
    const uint32_t ridx = uint32_t(int(t.idx(bx, by).x));
    const bool is_rendered = (x & 1) == offset_x(ridx) && (y & 1) == offset_y(ridx);
    st.wgt = out_of_screen ? 1.0f : is_rendered ? 0.2f : 0.0f;
}

__device__ inline ushort4 reconstruct_finish(const ReconstructArgs& a, const ReconstructStage& st) {
#pragma clang fp contract(off)
    float4 pre = a.previous ? reinhard(rgba16f_bilinear(a.previous, a.hw, a.hh, st.pu, st.pv)) : st.rendered;
    pre = min4(max4(pre, st.lo), st.hi);  // clamp(pre, aabb_min, aabb_max)
    const float wgt = st.wgt, iw = 1.0f - wgt;
    const float4 rd = st.rendered;
    const float4 r = make_float4(pre.x * iw + rd.x * wgt, pre.y * iw + rd.y * wgt, pre.z * iw + rd.z * wgt,
                                 pre.w * iw + rd.w * wgt);  // mix
    return store_rgba16f(inverse_reinhard(r));
}

__global__ __launch_bounds__(256) void reconstruct_kernel(ReconstructArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.hw * a.hh) return;
    const int x = i % a.hw, y = i / a.hw;
    const GlobalTiles t{a.index, a.render, a.hw / 2};
    ReconstructStage st;
    const float depth = a.cb[i];
    reconstruct_prepare(a, t, x, y, depth, linear_depth(a.lin, depth), a.distance[(y >> 1) * (a.hw / 2) + (x >> 1)],
                        st);
    a.out[i] = reconstruct_finish(a, st);
}

// The half-resolution inputs of the upscale from global memory at clamped coordinates.
struct GlobalHalf {
    const ushort4* reconstruct;
    const float* cb;
    int hw, hh;
    __device__ float4 rec(int hx, int hy) const {
        return load_rgba16f(reconstruct, clampi(hy, 0, hh - 1) * hw + clampi(hx, 0, hw - 1));
    }
    float lin_p[2];
    __device__ float lin(int hx, int hy) const {
        return linear_depth(lin_p, cb[clampi(hy, 0, hh - 1) * hw + clampi(hx, 0, hw - 1)]);
    }
};

// VolumetricCloudUpscale.comp for full-resolution pixel (x, y): `depth` its depth, `hdr` its
// RGBA16F colour; returns the composited colour.
template <class Half>
__device__ inline ushort4 upscale_pixel(const float lin_p[2], const Half& hs, int x, int y, float depth, ushort4 hdr) {
#pragma clang fp contract(off)
    const float lin = linear_depth(lin_p, depth);
    const int ox[4] = {-1, 1, 1, -1}, oy[4] = {1, 1, -1, -1};  // textureGather order
    float4 n[4], near_n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float min_delta = 1e10f;
    bool edge = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int hx = (x + ox[k]) >> 1, hy = (y + oy[k]) >> 1;
        n[k] = hs.rec(hx, hy);
        const float d = fabsf(lin - hs.lin(hx, hy));
        if (d < min_delta) {  // reconstructed_neighbors[nearest_i]
            near_n = n[k];
            min_delta = d;
        }
        if (d > lin * 0.1f) edge = true;
    }
    const float max_a = fmaxf(fmaxf(n[0].w, n[1].w), fmaxf(n[2].w, n[3].w));
    const float min_a = fminf(fminf(n[0].w, n[1].w), fminf(n[2].w, n[3].w));
    bool take_nearest = false;
    if (edge) take_nearest = (max_a - min_a) / fmaxf(1e-6f, 1.0f - min_a) > 0.2f;  // evaluated only at edges
    float4 up;
    if (take_nearest) {
        up = near_n;
    } else {
        up = make_float4(((n[0].x + n[1].x) + (n[2].x + n[3].x)) * 0.25f, ((n[0].y + n[1].y) + (n[2].y + n[3].y)) * 0.25f,
                         ((n[0].z + n[1].z) + (n[2].z + n[3].z)) * 0.25f, ((n[0].w + n[1].w) + (n[2].w + n[3].w)) * 0.25f);
    }
    float4 color = unpack_rgba16f(hdr);
    const float t = up.w <= kMinTransmittance ? 0.0f : up.w;
    color.x = color.x * t + up.x;
    color.y = color.y * t + up.y;
    color.z = color.z * t + up.z;
    return store_rgba16f(color);
}

struct UpscaleArgs {
    float lin[2];
    const float* cb;             // W/2 x H/2
    const float* depth;          // W x H
    const ushort4* reconstruct;  // W/2 x H/2
    ushort4* hdr;                // W x H, composited in place
    int w, h;
};

__global__ __launch_bounds__(256) void upscale_kernel(UpscaleArgs a) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.w * a.h) return;
    const GlobalHalf hs{a.reconstruct, a.cb, a.w / 2, a.h / 2, {a.lin[0], a.lin[1]}};
    a.hdr[i] = upscale_pixel(a.lin, hs, i % a.w, i / a.w, a.depth[i], a.hdr[i]);
}

// Reconstruction + upscale in one pass.  A workgroup owns a 16 x 16 tile of half-resolution
// pixels and the 32 x 32 full-resolution pixels above it:
//  1. the 12 x 12 quarter-resolution tiles its reconstruction reads ((index, linear depth) and
//     the Reinhard-mapped sample) go to LDS, mapped once per tile instead of per reader;
//  2. the reconstruction of the 18 x 18 half-resolution pixels (the tile and the 1-pixel ring
//     the upscale's diagonal neighbours reach) goes to LDS as RGBA16F, with the checkerboard
//     depth beside it; the tile's own pixels are also stored as the pass's output (the next
//     frame's history);
//  3. the upscale reads its 4 neighbours from LDS and composites the full-resolution pixels.
// LDS slot (l) holds the value at the clamped coordinate origin + l, so a lookup at an
// unclamped coordinate returns what the clamped global lookup returns.
constexpr int kRecTile = 16;
constexpr int kRecHalo = kRecTile + 2;
constexpr int kRecQuarter = kRecTile / 2 + 4;

struct LdsTiles {
    const float2 (*index)[kRecQuarter];
    const float4 (*color_)[kRecQuarter];
    int qx0, qy0;
    __device__ float2 idx(int qx, int qy) const { return index[qy - qy0][qx - qx0]; }
    __device__ float4 color(int qx, int qy) const { return color_[qy - qy0][qx - qx0]; }
};

struct LdsHalf {
    const ushort4 (*rec_)[kRecHalo];
    const float (*lin_)[kRecHalo];  // DepthToLinearDepth of the checkerboard depth
    int hx0, hy0;
    __device__ float4 rec(int hx, int hy) const { return unpack_rgba16f(rec_[hy - hy0][hx - hx0]); }
    __device__ float lin(int hx, int hy) const { return lin_[hy - hy0][hx - hx0]; }
};

struct ReconstructUpscaleArgs {
    ReconstructArgs r;
    const float* depth;  // W x H
    ushort4* hdr;        // W x H, composited in place
    int w, h;
};

__global__ __launch_bounds__(256) void reconstruct_upscale_kernel(ReconstructUpscaleArgs a) {
    __shared__ float4 s_color[kRecQuarter][kRecQuarter];
    __shared__ float2 s_index[kRecQuarter][kRecQuarter];
    __shared__ ushort4 s_rec[kRecHalo][kRecHalo];
    __shared__ float s_lin[kRecHalo][kRecHalo];
    const ReconstructArgs& r = a.r;
    const int hw = r.hw, hh = r.hh, qw = hw / 2, qh = hh / 2;
    const int hx0 = blockIdx.x * kRecTile - 1, hy0 = blockIdx.y * kRecTile - 1;
    const int qx0 = blockIdx.x * (kRecTile / 2) - 2, qy0 = blockIdx.y * (kRecTile / 2) - 2;
    for (int t = threadIdx.x; t < kRecQuarter * kRecQuarter; t += 256) {
        const int lx = t % kRecQuarter, ly = t / kRecQuarter;
        const int q = clampi(qy0 + ly, 0, qh - 1) * qw + clampi(qx0 + lx, 0, qw - 1);
        s_index[ly][lx] = r.index[q];
        s_color[ly][lx] = reinhard(load_rgba16f(r.render, q));
    }
    __syncthreads();  // phase 1: the quarter-resolution tiles
    const LdsTiles tiles{s_index, s_color, qx0, qy0};
    // Phase 2: the 18 x 18 reconstructions with their linear depths.
    for (int t = threadIdx.x; t < kRecHalo * kRecHalo; t += 256) {
        const int lx = t % kRecHalo, ly = t / kRecHalo;
        const int x = clampi(hx0 + lx, 0, hw - 1), y = clampi(hy0 + ly, 0, hh - 1);
        const float depth = r.cb[y * hw + x];
        const float lin = linear_depth(r.lin, depth);
        ReconstructStage st;
        reconstruct_prepare(r, tiles, x, y, depth, lin, r.distance[(y >> 1) * qw + (x >> 1)], st);
        const ushort4 v = reconstruct_finish(r, st);
        s_rec[ly][lx] = v;
        s_lin[ly][lx] = lin;
        if (hx0 + lx == x && hy0 + ly == y && lx >= 1 && lx <= kRecTile && ly >= 1 && ly <= kRecTile)
            r.out[y * hw + x] = v;
    }
    __syncthreads();
    // Phase 3: the upscale, the 4 pixels' depth and colour read before any is written.
    const LdsHalf hs{s_rec, s_lin, hx0, hy0};
    constexpr int kPixels = 4 * kRecTile * kRecTile / 256;
    int idx[kPixels], pxs[kPixels], pys[kPixels];
    bool inside[kPixels];
    float pdepth[kPixels];
    ushort4 hdr[kPixels];
#pragma unroll
    for (int k = 0; k < kPixels; ++k) {
        const int t = threadIdx.x + 256 * k;
        pxs[k] = blockIdx.x * 2 * kRecTile + t % (2 * kRecTile);
        pys[k] = blockIdx.y * 2 * kRecTile + t / (2 * kRecTile);
        inside[k] = pxs[k] < a.w && pys[k] < a.h;
        idx[k] = inside[k] ? pys[k] * a.w + pxs[k] : 0;
        pdepth[k] = a.depth[idx[k]];
        hdr[k] = a.hdr[idx[k]];
    }
#pragma unroll
    for (int k = 0; k < kPixels; ++k)
        if (inside[k]) a.hdr[idx[k]] = upscale_pixel(r.lin, hs, pxs[k], pys[k], pdepth[k], hdr[k]);
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

namespace {
int make_reconstruct_args(const SkyhipCloudCommonBufferData* common, const float* checkerboard,
                          const float* index_linear_depth, const void* render_rgba16f, const float* cloud_distance,
                          const void* previous_rgba16f, void* out_rgba16f, int width, int height, ReconstructArgs& a) {
    if (!common || !checkerboard || !index_linear_depth || !render_rgba16f || !cloud_distance || !out_rgba16f)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(check_viewport(width, height));
    if (!finite_block(common->uInvMVP, 16) || !finite_block(common->uReprojectMat, 16) ||
        !finite_block(common->uCameraPos, 3) || !finite_block(common->uLinearDepthParam, 2))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud common block has a non-finite field");
    if (previous_rgba16f == out_rgba16f) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "previous and output alias");
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
    return SKYHIP_OK;
}
}  // namespace

extern "C" int skyhip_cloud_reconstruct(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                        const float* checkerboard, const float* index_linear_depth,
                                        const void* render_rgba16f, const float* cloud_distance,
                                        const void* previous_rgba16f, void* out_rgba16f, int32_t width,
                                        int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    ReconstructArgs a;
    SKYHIP_CHECK(make_reconstruct_args(common, checkerboard, index_linear_depth, render_rgba16f, cloud_distance,
                                       previous_rgba16f, out_rgba16f, width, height, a));
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

extern "C" int skyhip_cloud_checkerboard_index(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                               const float* depth, int32_t width, int32_t height,
                                               float* out_checkerboard, float* out_index_linear_depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !depth || !out_checkerboard || !out_index_linear_depth)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(check_viewport(width, height));
    if (!finite_block(common->uLinearDepthParam, 2)) return fail(SKYHIP_ERR_NON_FINITE, "uLinearDepthParam");
    if ((reinterpret_cast<uintptr_t>(depth) & 15) || (reinterpret_cast<uintptr_t>(out_checkerboard) & 7) ||
        (reinterpret_cast<uintptr_t>(out_index_linear_depth) & 7))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "depth must be 16-byte aligned, the outputs 8-byte aligned");
    CheckerboardIndexArgs a;
    a.lin[0] = common->uLinearDepthParam[0];
    a.lin[1] = common->uLinearDepthParam[1];
    a.base = common->uBaseShadingIndex;
    a.depth = depth;
    a.w = width;
    a.h = height;
    a.cb = out_checkerboard;
    a.out = reinterpret_cast<float2*>(out_index_linear_depth);
    hipLaunchKernelGGL(checkerboard_index_kernel, dim3(div_up((long long)(width / 4) * (height / 4), 256)), dim3(256),
                       0, ctx->stream, a);
    SKYHIP_LAUNCH_CHECK("checkerboard_index_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_reconstruct_upscale(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                                const float* checkerboard, const float* index_linear_depth,
                                                const void* render_rgba16f, const float* cloud_distance,
                                                const void* previous_rgba16f, void* out_rgba16f, const float* depth,
                                                void* hdr_rgba16f, int32_t width, int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    ReconstructUpscaleArgs a;
    SKYHIP_CHECK(make_reconstruct_args(common, checkerboard, index_linear_depth, render_rgba16f, cloud_distance,
                                       previous_rgba16f, out_rgba16f, width, height, a.r));
    if (!depth || !hdr_rgba16f) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    a.depth = depth;
    a.hdr = reinterpret_cast<ushort4*>(hdr_rgba16f);
    a.w = width;
    a.h = height;
    const dim3 grid(div_up(a.r.hw, kRecTile), div_up(a.r.hh, kRecTile));
    hipLaunchKernelGGL(reconstruct_upscale_kernel, grid, dim3(256), 0, ctx->stream, a);
    SKYHIP_LAUNCH_CHECK("reconstruct_upscale_kernel");
    return SKYHIP_OK;
}
