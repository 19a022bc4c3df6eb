// Test infrastructure (the oracle): CPU restatement of the volumetric cloud's shadow passes,
// VolumetricCloud::RenderShadow (src/SkyRendering/VolumetricCloud.cpp:282-325):
// VolumetricCloudShadowMap.comp, VolumetricCloudShadowMapBlur.comp and
// VolumetricCloudShadowFroxel.comp with SampleCloudShadowTransmittance
// (VolumetricCloudShadowInterface.glsl:4-8).  The shadow-map sampler is LINEAR with
// CLAMP_TO_BORDER, border (1e10, 1) (VolumetricCloud.cpp:106-112); the froxel is R16 UNORM,
// stored as round(clamp(v, 0, 1) * 65535).  Only the tests and the benchmark's CPU baseline
// use this code.
#pragma once

#include "oracle_cloud.h"

namespace oracle {

// LINEAR / CLAMP_TO_BORDER(1e10, 1) lookup of an RG32F map.
template <class S>
void ShadowMapBilinear(const float* map, int w, int h, S u, S v, S out[2]) {
    S fx = gclamp(u * S(w) - S(0.5), S(-2), S(w) + S(1)), fy = gclamp(v * S(h) - S(0.5), S(-2), S(h) + S(1));
    S x0f = std::floor(fx), y0f = std::floor(fy);
    int x0 = int(x0f), y0 = int(y0f);
    S ax = fx - x0f, ay = fy - y0f;
    S c[4][2];
    for (int k = 0; k < 4; ++k) {
        int x = x0 + (k & 1), y = y0 + (k >> 1);
        bool border = x < 0 || y < 0 || x >= w || y >= h;
        c[k][0] = border ? S(1e10f) : S(map[2 * (size_t(y) * w + x)]);
        c[k][1] = border ? S(1) : S(map[2 * (size_t(y) * w + x) + 1]);
    }
    for (int ch = 0; ch < 2; ++ch)
        out[ch] = (c[0][ch] * (S(1) - ax) + c[1][ch] * ax) * (S(1) - ay) + (c[2][ch] * (S(1) - ax) + c[3][ch] * ax) * ay;
}

// VolumetricCloudShadowMap.comp: out[2 * i] = mid-shell depth, out[2 * i + 1] = transmittance.
template <class S>
void CloudShadowMap(const SkyhipCloudCommonBufferData& c, const SkyhipCloudMaterial& m, const uint16_t* bn,
                    int bn_size, const float* previous, int w, int h, S* out) {
    CloudCtx<S> C(&c, nullptr, &m);
    const V3<S> dir = -C.sun;
    const S rb = C.R + C.hb, rt = C.R + C.ht;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < w * h; ++i) {
        int x = i % w, y = i / w;
        S u = (S(x) + S(0.5)) / S(w), v = (S(y) + S(0.5)) / S(h);
        V3<S> origin = projective_mul<S>(c.uInvLightVP, {u * S(2) - S(1), v * S(2) - S(1), S(0)});
        V3<S> up = {origin.x, origin.y, origin.z + C.R};
        S r = length(up);
        up = up / r;
        S mu = dot(up, dir);
        S t1 = 0, t2 = 0;  // LineShellFirstIntersect (:19-35)
        S disc_top = r * r * (mu * mu - S(1)) + rt * rt;
        if (disc_top > S(0)) {
            S disc_bottom = r * r * (mu * mu - S(1)) + rb * rb;
            t1 = -r * mu - std::sqrt(disc_top);
            t2 = -r * mu + (disc_bottom >= S(0) ? -std::sqrt(disc_bottom) : std::sqrt(disc_top));
        }
        S dist = std::fmax(t2 - t1, S(0));
        dist = std::fmin(dist, S(1) / std::cos(S(85) * S(kPi) / S(180)) * (C.ht - C.hb));
        S od = 0;
        if (dist > S(0)) {
            // steps = mix(12, 6, |dir.z|) and its uint() in the shader's FP32
            float az = std::fabs(float(dir.z));
            float steps_f = 12.0f * (1.0f - az) + 6.0f * az;
            S steps = S(steps_f);
            S step = dist / steps;
            S noise = S(bn[size_t(y & 63) * bn_size + (x & 63)]) / S(65535);
            S jitter = noise + S(c.uFrameID) * S(0.61803398875);
            S t = t1 + step * (jitter - std::floor(jitter));
            for (uint32_t cnt = uint32_t(steps_f); cnt != 0; --cnt, t += step) {
                V3<S> pos = origin + dir * t;
                od += C.SampleSigmaT(pos, C.CalHeight01(pos)) * step;
            }
        }
        S res[2] = {t1 * S(0.5) + t2 * S(0.5), std::exp(-od)};
        if (previous) {
            V3<S> p = projective_mul<S>(c.uShadowMapReprojectMat, {u * S(2) - S(1), v * S(2) - S(1), S(0)});
            S pu = p.x * S(0.5) + S(0.5), pv = p.y * S(0.5) + S(0.5);
            if (pu >= S(0.5) / S(w) && pu <= (S(w) - S(0.5)) / S(w) && pv >= S(0.5) / S(h) &&
                pv <= (S(h) - S(0.5)) / S(h)) {
                S pre[2];
                ShadowMapBilinear<S>(previous, w, h, pu, pv, pre);
                for (int ch = 0; ch < 2; ++ch) res[ch] = pre[ch] * S(0.8) + res[ch] * S(0.2);
            }
        }
        out[2 * i] = res[0];
        out[2 * i + 1] = res[1];
    }
}

// One pass of VolumetricCloudShadowMapBlur.comp (axis 0: x, 1: y), in FP32.
inline void CloudShadowBlur(const float* in, float* out, int w, int h, int axis) {
    const float wt[5] = {0.227027f, 0.1945946f, 0.1216216f, 0.054054f, 0.016216f};
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int ch = 0; ch < 2; ++ch) {
                float r = in[2 * (size_t(y) * w + x) + ch] * wt[0];
                for (int k = 1; k < 5; ++k) {
                    size_t i1 = axis == 0 ? size_t(y) * w + std::max(x - k, 0) : size_t(std::max(y - k, 0)) * w + x;
                    size_t i2 = axis == 0 ? size_t(y) * w + std::min(x + k, w - 1)
                                          : size_t(std::min(y + k, h - 1)) * w + x;
                    r += in[2 * i1 + ch] * wt[k];
                    r += in[2 * i2 + ch] * wt[k];
                }
                out[2 * (size_t(y) * w + x) + ch] = r;
            }
}

// VolumetricCloudShadowFroxel.comp: out[(z * h + y) * w + x], R16 UNORM.
template <class S>
void CloudShadowFroxel(const SkyhipCloudCommonBufferData& c, const float* map, int mw, int mh, int w, int h, int d,
                       uint16_t* out) {
    const V3<S> cam = cvt3<S>(c.uCameraPos);
    const S step = S(c.uShadowFroxelMaxDistance) / S(d);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < w * h; ++i) {
        int x = i % w, y = i / w;
        S u = (S(x) + S(0.5)) / S(w), v = (S(y) + S(0.5)) / S(h);
        V3<S> frag = projective_mul<S>(c.uInvMVP, {u * S(2) - S(1), v * S(2) - S(1), S(0)});
        V3<S> dir = normalize(frag - cam);
        S sum = 0;
        for (int z = 0; z < d; ++z) {
            S t = (S(0.5) + S(z)) * step;
            V3<S> ndc = projective_mul<S>(c.uLightVP, cam + dir * t);
            S dt[2];
            ShadowMapBilinear<S>(map, mw, mh, ndc.x * S(0.5) + S(0.5), ndc.y * S(0.5) + S(0.5), dt);
            S s = gclamp((dt[0] - ndc.z) * S(2), S(0), S(1));
            sum += dt[1] * (S(1) - s) + s;
            S vis = gclamp(sum / S(z + 1), S(0), S(1));
            out[(size_t(z) * h + y) * w + x] = uint16_t(std::nearbyint(vis * S(65535)));
        }
    }
}

}  // namespace oracle
