// Test infrastructure (the oracle): CPU restatement of the volumetric cloud's checkerboard,
// shading-index, temporal-reconstruction and upscale passes, VolumetricCloud::Render
// (src/SkyRendering/VolumetricCloud.cpp:327-422): CheckerboardGen.comp,
// VolumetricCloudIndexGen.comp, VolumetricCloudReconstruct.comp and
// VolumetricCloudUpscale.comp.  Texture semantics: textureGather and texelFetchClamp on the
// NEAREST / CLAMP_TO_EDGE samplers the reference binds (:337, :351, :394-398, :411-413) read
// the clamped texel; an out-of-range texelFetch (the index pass's neighbour tiles) reads 0;
// the previous reconstruction is sampled LINEAR / CLAMP_TO_EDGE; RGBA16F stores round to
// nearest even.  In FP32 (S = float, no FMA contraction) each function evaluates the
// shaders' expressions in their order.  Only the tests and the benchmark's CPU baseline use
// this code.
#pragma once

#include "oracle_cloud.h"
#include "oracle_pathtrace.h"

namespace oracle {

namespace recon {

template <class S>
S LinearDepth(const float p[2], S d) {  // DepthToLinearDepth, VolumetricCloudCommon.glsl:32-34
    return S(1) / (S(p[0]) - S(p[1]) * d);
}
inline int OffX(uint32_t i) { return int(((i + 1u) >> 1) & 1u); }  // IndexToOffset, :42-52
inline int OffY(uint32_t i) { return int(((i + 2u) >> 1) & 1u); }
inline int Clampi(int v, int lo, int hi) { return std::min(std::max(v, lo), hi); }

template <class S>
struct C4 {
    S v[4];
};
template <class S>
C4<S> LoadHalf4(const uint16_t* p, size_t i) {
    return {{S(half_to_float(p[4 * i])), S(half_to_float(p[4 * i + 1])), S(half_to_float(p[4 * i + 2])),
             S(half_to_float(p[4 * i + 3]))}};
}
template <class S>
C4<S> Reinhard(C4<S> c) {
    for (int k = 0; k < 3; ++k) c.v[k] = c.v[k] / (S(1) + c.v[k]);
    return c;
}
template <class S>
C4<S> InverseReinhard(C4<S> c) {
    for (int k = 0; k < 3; ++k) c.v[k] = c.v[k] / (S(1) - c.v[k]);
    return c;
}
template <class S>
V3<S> ProjectiveMul(const float* m, V3<S> v) {  // Common.glsl:7-10, in the shader's order
    S x = S(m[0]) * v.x + S(m[4]) * v.y + S(m[8]) * v.z + S(m[12]);
    S y = S(m[1]) * v.x + S(m[5]) * v.y + S(m[9]) * v.z + S(m[13]);
    S z = S(m[2]) * v.x + S(m[6]) * v.y + S(m[10]) * v.z + S(m[14]);
    S w = S(m[3]) * v.x + S(m[7]) * v.y + S(m[11]) * v.z + S(m[15]);
    return {x / w, y / w, z / w};
}

}  // namespace recon

// CheckerboardGen.comp: out[y * w/2 + x] = max of the 2x2 block (2x, 2y) of depth where x, y
// have equal parity, else min.
inline void CheckerboardDepth(const float* depth, int w, int h, float* out) {
    const int hw = w / 2, hh = h / 2;
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < hh; ++y)
        for (int x = 0; x < hw; ++x) {
            const float a = depth[size_t(2 * y) * w + 2 * x], b = depth[size_t(2 * y) * w + 2 * x + 1];
            const float c = depth[size_t(2 * y + 1) * w + 2 * x], d = depth[size_t(2 * y + 1) * w + 2 * x + 1];
            const bool bmax = (x & 1) == (y & 1);
            out[size_t(y) * hw + x] = bmax ? std::max(std::max(a, b), std::max(c, d))
                                           : std::min(std::min(a, b), std::min(c, d));
        }
}

// VolumetricCloudIndexGen.comp: out[2 * i] = index, out[2 * i + 1] = its linear depth, over
// the (cw/2) x (ch/2) tiles of the checkerboard cb (cw x ch).
template <class S>
void ShadingIndex(const SkyhipCloudCommonBufferData& c, const float* cb, int cw, int ch, S* out) {
    using namespace recon;
    const int qw = cw / 2, qh = ch / 2;
    const uint32_t base = c.uBaseShadingIndex;
    auto pos_to_index = [&](int x, int y) { return (base + uint32_t((x + y) & 1)) & 3u; };
    const int ox[8] = {0, 1, 1, 1, 0, -1, -1, -1}, oy[8] = {1, 1, 0, -1, -1, -1, 0, 1};
    #pragma omp parallel for schedule(static)
    for (int py = 0; py < qh; ++py)
        for (int px = 0; px < qw; ++px) {
            float tile[4];
            for (uint32_t k = 0; k < 4; ++k) tile[k] = cb[size_t(2 * py + OffY(k)) * cw + 2 * px + OffX(k)];
            uint32_t nearest = 0, farthest = 0;
            for (uint32_t k = 1; k < 4; ++k) {
                nearest = tile[k] < tile[nearest] ? k : nearest;
                farthest = tile[k] > tile[farthest] ? k : farthest;
            }
            const S near_lin = LinearDepth<S>(c.uLinearDepthParam, S(tile[nearest]));
            const S far_lin = LinearDepth<S>(c.uLinearDepthParam, S(tile[farthest]));
            int close_near = 0, close_far = 0;
            for (int k = 0; k < 8; ++k) {
                const int tx = px + ox[k], ty = py + oy[k];
                const uint32_t idx = pos_to_index(tx, ty);
                const int x = 2 * tx + OffX(idx), y = 2 * ty + OffY(idx);
                const float d = (x < 0 || y < 0 || x >= cw || y >= ch) ? 0.0f : cb[size_t(y) * cw + x];
                const S lin = LinearDepth<S>(c.uLinearDepthParam, S(d));
                const S max_delta = lin * S(0.25);
                if (std::fabs(lin - near_lin) < max_delta) ++close_near;
                if (std::fabs(lin - far_lin) < max_delta) ++close_far;
            }
            const uint32_t index = close_near == 0 ? nearest : close_far == 0 ? farthest : pos_to_index(px, py);
            out[2 * (size_t(py) * qw + px)] = S(index);
            out[2 * (size_t(py) * qw + px) + 1] = LinearDepth<S>(c.uLinearDepthParam, S(tile[index]));
        }
}

// VolumetricCloudReconstruct.comp (USE_TONEMAPPING 1, USE_VARIANCE 0) over the (w/2) x (h/2)
// half-resolution pixels; index (RG32F), render (RGBA16F) and distance (R32F) are
// (w/4) x (h/4); previous (RGBA16F half resolution) may be null: every pixel then takes the
// new sample.  out: 4 values per pixel, before the RGBA16F store.
template <class S>
void Reconstruct(const SkyhipCloudCommonBufferData& c, const float* cb, const float* index, const uint16_t* render,
                 const float* distance, const uint16_t* previous, int w, int h, S* out) {
    using namespace recon;
    const int hw = w / 2, hh = h / 2, qw = w / 4, qh = h / 4;
    const int ox[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1}, oy[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
    const V3<S> cam = cvt3<S>(c.uCameraPos);
    auto qidx = [&](int x, int y) { return size_t(Clampi(y, 0, qh - 1)) * qw + Clampi(x, 0, qw - 1); };
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < hh; ++y)
        for (int x = 0; x < hw; ++x) {
            const int bx = x >> 1, by = y >> 1;
            const S u = (S(x) + S(0.5)) / S(hw), v = (S(y) + S(0.5)) / S(hh);
            const float depth = cb[size_t(y) * hw + x];
            const S lin = LinearDepth<S>(c.uLinearDepthParam, S(depth));
            S rl[9], delta[9], min_delta = S(1e10f);
            int nearest = 0;
            for (int k = 0; k < 9; ++k) {
                rl[k] = S(index[2 * qidx(bx + ox[k], by + oy[k]) + 1]);
                delta[k] = std::fabs(rl[k] - lin);
                if (delta[k] < min_delta) {
                    min_delta = delta[k];
                    nearest = k;
                }
            }
            const C4<S> nc = Reinhard(LoadHalf4<S>(render, qidx(bx + ox[nearest], by + oy[nearest])));
            C4<S> lo = nc, hi = nc;
            for (int k = 0; k < 9; ++k) {
                const C4<S> r = Reinhard(LoadHalf4<S>(render, qidx(bx + ox[k], by + oy[k])));
                if (delta[k] < rl[k] * S(0.3) ||
                    std::fabs(r.v[3] - nc.v[3]) / std::max(S(1e-6f), S(1) - std::max(r.v[3], nc.v[3])) < S(0.2))
                    for (int ch = 0; ch < 4; ++ch) {
                        lo.v[ch] = std::min(lo.v[ch], r.v[ch]);
                        hi.v[ch] = std::max(hi.v[ch], r.v[ch]);
                    }
            }
            const C4<S> rendered = Reinhard(LoadHalf4<S>(render, size_t(by) * qw + bx));
            const V3<S> frag = ProjectiveMul<S>(c.uInvMVP, {u * S(2) - S(1), v * S(2) - S(1), S(depth) * S(2) - S(1)});
            V3<S> dir = frag - cam;
            dir = dir * (S(1) / std::sqrt(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z));
            const V3<S> cloud_pos = cam + dir * S(distance[size_t(by) * qw + bx]);
            const V3<S> pre_ndc = ProjectiveMul<S>(c.uReprojectMat, cloud_pos);
            const S pu = pre_ndc.x * S(0.5) + S(0.5), pv = pre_ndc.y * S(0.5) + S(0.5);
            const bool out_of_screen = !previous || std::max(std::fabs(pre_ndc.x), std::fabs(pre_ndc.y)) > S(1);
            C4<S> pre = rendered;
            if (previous) {  // LINEAR / CLAMP_TO_EDGE
                S fx = pu * S(hw) - S(0.5), fy = pv * S(hh) - S(0.5);
                S x0f = std::floor(fx), y0f = std::floor(fy);
                S ax = fx - x0f, ay = fy - y0f;
                int x0 = int(std::fmin(std::fmax(x0f, S(-1)), S(hw))), y0 = int(std::fmin(std::fmax(y0f, S(-1)), S(hh)));
                const int xa = Clampi(x0, 0, hw - 1), xb = Clampi(x0 + 1, 0, hw - 1);
                const int ya = Clampi(y0, 0, hh - 1), yb = Clampi(y0 + 1, 0, hh - 1);
                const C4<S> c00 = LoadHalf4<S>(previous, size_t(ya) * hw + xa);
                const C4<S> c10 = LoadHalf4<S>(previous, size_t(ya) * hw + xb);
                const C4<S> c01 = LoadHalf4<S>(previous, size_t(yb) * hw + xa);
                const C4<S> c11 = LoadHalf4<S>(previous, size_t(yb) * hw + xb);
                const S bx1 = S(1) - ax, by1 = S(1) - ay;
                for (int ch = 0; ch < 4; ++ch)
                    pre.v[ch] = (c00.v[ch] * bx1 + c10.v[ch] * ax) * by1 + (c01.v[ch] * bx1 + c11.v[ch] * ax) * ay;
                pre = Reinhard(pre);
            }
            for (int ch = 0; ch < 4; ++ch) pre.v[ch] = std::min(std::max(pre.v[ch], lo.v[ch]), hi.v[ch]);
            const uint32_t ridx = uint32_t(int(index[2 * (size_t(by) * qw + bx)]));
            const bool is_rendered = (x & 1) == OffX(ridx) && (y & 1) == OffY(ridx);
            const S wgt = out_of_screen ? S(1) : is_rendered ? S(0.2f) : S(0);
            C4<S> r;
            for (int ch = 0; ch < 4; ++ch) r.v[ch] = pre.v[ch] * (S(1) - wgt) + rendered.v[ch] * wgt;
            r = InverseReinhard(r);
            for (int ch = 0; ch < 4; ++ch) out[4 * (size_t(y) * hw + x) + ch] = r.v[ch];
        }
}

// VolumetricCloudUpscale.comp: out (4 per full-resolution pixel, before the RGBA16F store) =
// hdr's colour composited under the upscaled reconstruction.
template <class S>
void Upscale(const SkyhipCloudCommonBufferData& c, const float* cb, const float* depth, const uint16_t* reconstruct,
             const uint16_t* hdr, int w, int h, S* out) {
    using namespace recon;
    const int hw = w / 2, hh = h / 2;
    const int ox[4] = {-1, 1, 1, -1}, oy[4] = {1, 1, -1, -1};
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const S lin = LinearDepth<S>(c.uLinearDepthParam, S(depth[size_t(y) * w + x]));
            C4<S> n[4];
            S min_delta = S(1e10f);
            int nearest = 0;
            bool edge = false;
            for (int k = 0; k < 4; ++k) {  // (p + o) >> 1 is an arithmetic shift
                const size_t q = size_t(Clampi((y + oy[k]) >> 1, 0, hh - 1)) * hw + Clampi((x + ox[k]) >> 1, 0, hw - 1);
                n[k] = LoadHalf4<S>(reconstruct, q);
                const S d = std::fabs(lin - LinearDepth<S>(c.uLinearDepthParam, S(cb[q])));
                if (d < min_delta) {
                    nearest = k;
                    min_delta = d;
                }
                if (d > lin * S(0.1f)) edge = true;
            }
            const S max_a = std::max(std::max(n[0].v[3], n[1].v[3]), std::max(n[2].v[3], n[3].v[3]));
            const S min_a = std::min(std::min(n[0].v[3], n[1].v[3]), std::min(n[2].v[3], n[3].v[3]));
            C4<S> up;
            if (edge && (max_a - min_a) / std::max(S(1e-6f), S(1) - min_a) > S(0.2f)) {
                up = n[nearest];
            } else {
                for (int ch = 0; ch < 4; ++ch)
                    up.v[ch] = ((n[0].v[ch] + n[1].v[ch]) + (n[2].v[ch] + n[3].v[ch])) * S(0.25);
            }
            const C4<S> color = LoadHalf4<S>(hdr, size_t(y) * w + x);
            const S t = up.v[3] <= S(0.01f) ? S(0) : up.v[3];
            S* o = out + 4 * (size_t(y) * w + x);
            for (int ch = 0; ch < 3; ++ch) o[ch] = color.v[ch] * t + up.v[ch];
            o[3] = color.v[3];
        }
}

}  // namespace oracle
