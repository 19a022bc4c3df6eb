// Test infrastructure (the oracle): CPU restatement of the volumetric-cloud
// raymarch, shaders/SkyRendering/VolumetricCloudRender.comp,
// VolumetricCloudCommon.glsl, VolumetricCloudShadowInterface.glsl:10-13 and the
// material plugins VolumetricCloudDefaultMaterial{Common,0,1}.glsl,
// VolumetricCloudMaterialMinimal.glsl and VolumetricCloudMaterialVoxel.glsl.
// The cloud shadow froxel is the constant 1 (the shadow passes are not part
// of this path); shadow-loop counters stay FP32 as in the shader.
#pragma once

#include "oracle_atmosphere.h"

namespace oracle {

struct CloudTextures {
    const SkyhipCloudMaterial* m;
};

// The raymarch's dual-lobe phase, set by oracle_bind_cloud_phase as the library's
// skyhip_bind_cloud_phase sets a context's; the reference's constant lobes by default
// (VolumetricCloudRender.comp:125-126).
inline SkyhipCloudPhase ReferencePhase() { return SkyhipCloudPhase{-0.15f, 2.16f, 0.85f, 0.0f}; }
inline SkyhipCloudPhase& BoundPhase() {
    static SkyhipCloudPhase p = ReferencePhase();
    return p;
}

template <class S>
struct CloudCtx {
    const SkyhipCloudCommonBufferData* c;
    const SkyhipCloudBufferData* b;
    const SkyhipCloudMaterial* m;
    V3<S> cam, sun;
    S R, hb, ht;

    CloudCtx(const SkyhipCloudCommonBufferData* c_, const SkyhipCloudBufferData* b_, const SkyhipCloudMaterial* m_)
        : c(c_), b(b_), m(m_) {
        cam = cvt3<S>(c->uCameraPos);
        sun = cvt3<S>(c->uSunDirection);
        R = c->uEarthRadius;
        hb = c->uBottomAltitude;
        ht = c->uTopAltitude;
    }

    // VolumetricCloudCommon.glsl:36-40
    S CalHeight01(V3<S> pos) const {
        S altitude = length(V3<S>{pos.x, pos.y, pos.z + R}) - R;
        return gclamp((altitude - hb) / (ht - hb), S(0), S(1));
    }

    // VolumetricCloudDefaultMaterialCommon.glsl:20-24
    void GetUVWLod(V3<S> pos, const SkyhipSampleInfo& si, S uvw[3], S& lod) const {
        uvw[0] = pos.x * S(si.frequency) + S(si.bias[0]);
        uvw[1] = pos.y * S(si.frequency) + S(si.bias[1]);
        uvw[2] = pos.z * S(si.frequency);
        lod = std::log2(S(si.k_lod) * distance(pos, cam)) + S(m->common.uLodBias);
    }

    S SampleDetail(V3<S> pos) const {
        S uvw[3], lod;
        GetUVWLod(pos, m->common.uDetailSampleInfo, uvw, lod);
        int dim[3] = {m->detail_size, m->detail_size, m->detail_size};
        return texlod3d_u8(m->detail, dim, m->detail_levels, uvw[0], uvw[1], uvw[2], lod, false);
    }

    // The Worley detail volume of SKYHIP_MATERIAL_DEFAULT1_WORLEY, sampled as the detail volume is.
    S SampleWorley(V3<S> pos) const {
        S uvw[3], lod;
        GetUVWLod(pos, m->worley.uWorleySampleInfo, uvw, lod);
        int dim[3] = {m->worley_size, m->worley_size, m->worley_size};
        return texlod3d_u8(m->worley_texture, dim, m->worley_levels, uvw[0], uvw[1], uvw[2], lod, false);
    }

    // Material SampleSigmaT(pos, height01)
    S SampleSigmaT(V3<S> pos, S height01) const {
        switch (m->kind) {
            case SKYHIP_MATERIAL_DEFAULT1:  // VolumetricCloudDefaultMaterial1.glsl:14-28
            case SKYHIP_MATERIAL_DEFAULT1_WORLEY: {
                const SkyhipCloudMaterial1BufferData& p = m->material1;
                S uvw[3], lod, ct[2];
                GetUVWLod(pos, m->common.uCloudMapSampleInfo, uvw, lod);
                texlod2d_u8_repeat(m->cloud_map, m->cloud_map_size, m->cloud_map_levels, 2, uvw[0], uvw[1], lod, 2, ct);
                S density = gclamp((ct[0] - S(p.uBaseDensityThreshold)) * S(p.uBaseEdgeHardness), S(0), S(1));
                density *= gclamp((S(1) - height01) * S(p.uBaseHeightHardness), S(0), S(1));
                if (density == S(0)) return S(0);
                S detail = SampleDetail(pos);
                S mask = std::fmax(gclamp(height01 - S(p.uHeightCut), S(0), S(1)),
                                   gclamp(S(p.uEdgeCur) - ct[0], S(0), S(1)));
                detail = (detail + S(p.uDetailBase)) * S(p.uDetailScale);
                detail *= mask;
                density = gclamp(density - detail, S(0), S(1));
                if (m->kind == SKYHIP_MATERIAL_DEFAULT1_WORLEY) {
                    // The second erosion (SkyhipCloudMaterialKind, include/skyhip.h).
                    S w = (SampleWorley(pos) + S(m->worley.uWorleyBase)) * S(m->worley.uWorleyScale);
                    density = gclamp(density - w * mask, S(0), S(1));
                }
                return density * S(m->common.uDensity) * height01;
            }
            case SKYHIP_MATERIAL_DEFAULT0: {  // VolumetricCloudDefaultMaterial0.glsl:18-31
                const SkyhipCloudMaterial0BufferData& p = m->material0;
                S uvw[3], lod, ct[2], d0[4], d1[4];
                GetUVWLod(pos, m->common.uCloudMapSampleInfo, uvw, lod);
                texlod2d_u8_repeat(m->cloud_map, m->cloud_map_size, m->cloud_map_levels, 2, uvw[0], uvw[1], lod, 2, ct);
                GetUVWLod(pos, m->common.uDisplacementSampleInfo, uvw, lod);
                texlod2d_u8_repeat(m->displacement, m->displacement_size, m->displacement_levels, 4, uvw[0], uvw[1],
                                   lod, 4, d0);
                texlod2d_u8_repeat(m->displacement, m->displacement_size, m->displacement_levels, 4, uvw[0], uvw[2],
                                   lod, 4, d1);
                V3<S> disp = {d0[0] + d1[2], d0[1], d1[3]};
                pos = pos + S(p.uDisplacementScale) * disp;
                S detail = SampleDetail(pos);
                detail = detail * S(p.uDetailParam[0]) + S(p.uDetailParam[1]);
                // CalHeightMask (:9-12) and Remap01 (:14-16)
                S hit = gclamp(height01 / ct[1], S(0), S(1));
                S mask = gclamp(hit * (hit - S(1)) * S(-4), S(0), S(1));
                S x = ct[0] * mask;
                return gclamp((x - detail) / (S(1) - detail), S(0), S(1)) * height01 * S(m->common.uDensity);
            }
            case SKYHIP_MATERIAL_MINIMAL:  // VolumetricCloudMaterialMinimal.glsl:6-7
                return S(m->minimal.uDensity);
            case SKYHIP_MATERIAL_VOXEL: {  // VolumetricCloudMaterialVoxel.glsl:12-16
                const SkyhipCloudVoxelBufferData& p = m->voxel;
                S u = pos.x * S(p.uSampleFrequency[0]) + S(p.uSampleBias[0]);
                S v = pos.y * S(p.uSampleFrequency[1]) + S(p.uSampleBias[1]);
                S lod = std::log2(S(p.uSampleLodK) * distance(pos, cam)) + S(p.uLodBias);
                return texlod3d_u8(m->voxel_texture, m->voxel_dim, m->voxel_levels, u, v, height01, lod, true) *
                       S(p.uDensity);
            }
        }
        return S(0);
    }

    // VolumetricCloudRender.comp:98-114 (float counter, t^2 distribution)
    S SampleShadow(V3<S> pos) const {
        S optical_depth = 0;
        float inv_shadow_steps = 1.0f / b->uShadowSteps;
        V3<S> sample_vector = S(b->uShadowDistance) * sun;
        float previous_t = 0.0f;
        for (float t = inv_shadow_steps; t <= 1.0f; t += inv_shadow_steps) {
            float current_t = t * t;
            float delta_t = current_t - previous_t;
            V3<S> sample_pos = pos + sample_vector * (S(previous_t) + S(0.5) * S(delta_t));
            S h01 = CalHeight01(sample_pos);
            optical_depth += SampleSigmaT(sample_pos, h01) * S(b->uShadowDistance) * S(delta_t);
            previous_t = current_t;
        }
        return std::exp(-optical_depth);
    }

    // VolumetricCloudCommon.glsl:58-63
    static S HenyeyGreenstein(S cos_theta, S g) {
        S a = S(1) - g * g;
        S bb = S(1) + g * g - S(2) * g * cos_theta;
        bb *= std::sqrt(bb);
        return (S(0.25) * S(1.0 / kPi)) * a / bb;
    }
};

template <class S>
struct MarchState {
    S t;
    V3<S> pos;
    S height01, step_size, transmittance, transmittance_sum, weighted_t_sum;
    S sun_env[2];
    S cos_sun_view;
};

// VolumetricCloudRender.comp:116-137
template <class S>
void RayMarchStep(const CloudCtx<S>& C, MarchState<S>& ctx) {
    S sigma_t = C.SampleSigmaT(ctx.pos, ctx.height01);
    if (sigma_t < S(1e-5)) return;
    S tr = std::exp(-ctx.step_size * sigma_t);
    S transmittance_to_sun = C.SampleShadow(ctx.pos);
    // The lobes are the reference's constants unless oracle_bind_cloud_phase set others.
    const SkyhipCloudPhase& ph = BoundPhase();
    S phase = gmix(CloudCtx<S>::HenyeyGreenstein(ctx.cos_sun_view, S(ph.back_g)) * S(ph.back_scale),
                   CloudCtx<S>::HenyeyGreenstein(ctx.cos_sun_view, S(ph.forward_g)),
                   std::exp(-S(C.b->uSunMultiscatteringSigmaScale) * sigma_t));
    S se[2];
    se[0] = transmittance_to_sun * phase;
    se[1] = gmix(S(C.b->uEnvBottomVisibility), S(1), ctx.height01);
    se[1] = se[1] - se[1] * std::exp(-S(C.b->uEnvMultiscatteringSigmaScale) * sigma_t);
    for (int k = 0; k < 2; ++k) {
        se[k] = se[k] - se[k] * tr;
        ctx.sun_env[k] += ctx.transmittance * se[k];
    }
    ctx.transmittance_sum += ctx.transmittance;
    ctx.weighted_t_sum += ctx.t * ctx.transmittance;
    ctx.transmittance *= tr;
}

// Pixel inputs of the render pass.
struct CloudImages {
    const float* transmittance;  // RGBA32F T LUT
    int tw, th;
    const float* ap_luminance;  // RGBA32F 3D
    const float* ap_transmittance;
    int aw, ah, ad;
    const uint16_t* blue_noise;
    int bn_size;
    const float* depth;  // R32F at (image_w x image_h) or null (= 1, far plane)
    const float* index;  // RG32F at (width x height) shading index (checkerboard mode) or null
    int image_w, image_h;
};

// The ray-scatter-visibility froxel bound for the raymarch and the path tracer (R16 UNORM
// [d][h][w]), or null: visibility 1.  Set by oracle_bind_shadow_froxel, as the library's
// skyhip_bind_shadow_froxel binds it to a context.
struct FroxelBinding {
    const uint16_t* data = nullptr;
    int w = 0, h = 0, d = 0;
};
inline FroxelBinding& BoundFroxel() {
    static FroxelBinding f;
    return f;
}

// SampleRayScatterVisibility (VolumetricCloudShadowInterface.glsl:10-13): texture(froxel,
// vec3(uv, dist * inv_max)) with LINEAR / CLAMP_TO_EDGE (GetShadowFroxel, VolumetricCloud.h:71-73)
// on R16 UNORM texels, then mix(1, tex, clamp(1 / w, 0, 1)).
template <class S>
S RayScatterVisibility(const FroxelBinding& f, S u, S v, S dist, S inv_max) {
    S w = dist * inv_max;
    S c[3] = {gclamp(u * S(f.w) - S(0.5), S(0), S(f.w - 1)), gclamp(v * S(f.h) - S(0.5), S(0), S(f.h - 1)),
              gclamp(w * S(f.d) - S(0.5), S(0), S(f.d - 1))};
    int n[3] = {f.w, f.h, f.d}, i0[3], i1[3];
    S a[3];
    for (int k = 0; k < 3; ++k) {
        i0[k] = int(std::floor(c[k]));
        i1[k] = std::min(i0[k] + 1, n[k] - 1);
        a[k] = c[k] - S(i0[k]);
    }
    S tex = 0;
    for (int k = 0; k < 8; ++k) {
        int x = (k & 1) ? i1[0] : i0[0], y = (k & 2) ? i1[1] : i0[1], z = (k & 4) ? i1[2] : i0[2];
        S wgt = ((k & 1) ? a[0] : S(1) - a[0]) * ((k & 2) ? a[1] : S(1) - a[1]) * ((k & 4) ? a[2] : S(1) - a[2]);
        tex += S(f.data[(size_t(z) * f.h + y) * f.w + x]) / S(65535) * wgt;
    }
    return gmix(S(1), tex, gclamp(S(1) / w, S(0), S(1)));
}

template <class S>
struct CloudPixelOut {
    S rgba[4];
    S distance;
    int steps;  // executed primary steps
};

// VolumetricCloudRender.comp:139-210 for output pixel (px, py).
//
// Checker mode: `forced` (non-null) holds another implementation's discrete
// decisions for this pixel: forced[0] the executed steps, forced[1] the scheduled
// steps uint(max(N dist / dist_max, 1)), each segment 0 in bits 0-15 and
// segment 1 in bits 16-31.  The march then uses those step counts (step size
// dist / scheduled) instead of deciding them itself, and reports how far the
// replayed decisions are from this evaluation's own: *margin = the largest
// |T - 0.01| / 0.01 over early exits it would have taken the other way, and
// *step_margin = the largest |x - round(x)| of x = N dist / dist_max over
// segments whose scheduled count differs (0 when every decision agrees).
// `trace` (optional): per executed step, 8 values (t, x, y, z, height01, sigma_t,
// transmittance after the step, step size), at most trace_max steps.
template <class S>
CloudPixelOut<S> CloudRaymarchPixel(const Atmosphere<S>& A, const CloudCtx<S>& C, const CloudImages& im, int px,
                                    int py, const uint32_t* forced = nullptr, S* margin = nullptr,
                                    double* trace = nullptr, int trace_max = 0, S* step_margin = nullptr) {
    const SkyhipCloudCommonBufferData& cb = *C.c;
    const SkyhipCloudBufferData& b = *C.b;
    int pcx = px, pcy = py;
    if (im.index) {
        uint32_t index = uint32_t(im.index[(size_t(py) * (im.image_w / 2) + px) * 2]);
        pcx = px * 2 + int(((index + 1) >> 1) & 1);  // IndexToOffset, VolumetricCloudCommon.glsl:42-52
        pcy = py * 2 + int(((index + 2) >> 1) & 1);
    }
    S u = (S(pcx) + S(0.5)) / S(im.image_w), v = (S(pcy) + S(0.5)) / S(im.image_h);
    S depth = im.depth ? S(im.depth[size_t(pcy) * im.image_w + pcx]) : S(1);
    V3<S> frag_pos = projective_mul<S>(cb.uInvMVP, {u * S(2) - S(1), v * S(2) - S(1), depth * S(2) - S(1)});
    V3<S> view_dir = normalize(frag_pos - C.cam);

    S r = C.cam.z + C.R;
    S mu = view_dir.z;
    // RayShellIntersect (:39-71)
    S t1[2] = {0, 0}, t2[2] = {0, 0};
    {
        S bottom_radius = C.R + C.hb, top_radius = C.R + C.ht;
        S db = r * r * (mu * mu - S(1)) + bottom_radius * bottom_radius;
        S dt = r * r * (mu * mu - S(1)) + top_radius * top_radius;
        S sb = std::sqrt(db), st = std::sqrt(dt);
        if (C.cam.z < C.hb) {
            t1[0] = -r * mu + sb;
            t2[0] = -r * mu + st;
        } else if (C.cam.z < C.ht) {
            if (db >= S(0) && mu < S(0)) {
                t2[0] = -r * mu - sb;
                t1[1] = -r * mu + sb;
                t2[1] = -r * mu + st;
            } else {
                t2[0] = -r * mu + st;
            }
        } else {
            if (db >= S(0) && mu < S(0)) {
                t1[0] = -r * mu - st;
                t2[0] = -r * mu - sb;
                t1[1] = -r * mu + sb;
                t2[1] = -r * mu + st;
            } else if (dt >= S(0) && mu < S(0)) {
                t1[0] = -r * mu - st;
                t2[0] = -r * mu + st;
            }
        }
    }
    S frag_dist = distance(frag_pos, C.cam);
    for (int i = 0; i < 2; ++i) t2[i] = gclamp(std::fmin(frag_dist, S(b.uMaxVisibleDistance)), t1[i], t2[i]);

    MarchState<S> ctx;
    ctx.cos_sun_view = dot(C.sun, view_dir);
    S dist = t2[0] - t1[0];
    dist = std::fmin(dist, S(b.uMaxRaymarchDistance));
    // uint(max(x, 1.0)): truncation toward zero
    uint32_t num_steps = 0;  // scheduled below (needs the segment lambdas)
    ctx.transmittance = 1;
    ctx.transmittance_sum = 0;
    ctx.weighted_t_sum = 0;
    ctx.sun_env[0] = ctx.sun_env[1] = 0;
    S noise = S(im.blue_noise[(py & (im.bn_size - 1)) * im.bn_size + (px & (im.bn_size - 1))]) / S(65535);
    S jitter = fract(noise + S(cb.uFrameID) * S(0.61803398875));
    int steps = 0;
    S worst = 0;
    // One march segment: the shader's loop, or the forced count in checker mode.
    S worst_steps = 0;
    // Scheduled step count of a segment; in checker mode the replayed one.
    auto schedule = [&](S seg_dist, int seg) -> uint32_t {
        S x = S(b.uMaxRaymarchSteps) * (seg_dist / S(b.uMaxRaymarchDistance));
        uint32_t own = uint32_t(std::fmax(x, S(1)));  // uint(max(x, 1.0)): truncation
        if (!forced) return own;
        uint32_t replay = (forced[1] >> (16 * seg)) & 0xFFFFu;
        if (replay != own) worst_steps = std::fmax(worst_steps, std::fabs(x - std::round(x)));
        return replay == 0 ? own : replay;
    };
    auto segment = [&](uint32_t n_steps, int seg) {
        uint32_t forced_n = forced ? ((forced[0] >> (16 * seg)) & 0xFFFFu) : 0u;
        uint32_t k = 0;
        for (uint32_t cnt = n_steps; cnt != 0; cnt--, ctx.t += ctx.step_size) {
            ctx.pos = C.cam + view_dir * ctx.t;  // UpdateContext (:90-93)
            ctx.height01 = C.CalHeight01(ctx.pos);
            RayMarchStep(C, ctx);
            if (trace && steps < trace_max) {
                double* tr = trace + 8 * steps;
                tr[0] = double(ctx.t), tr[1] = double(ctx.pos.x), tr[2] = double(ctx.pos.y), tr[3] = double(ctx.pos.z);
                tr[4] = double(ctx.height01), tr[5] = double(C.SampleSigmaT(ctx.pos, ctx.height01));
                tr[6] = double(ctx.transmittance), tr[7] = double(ctx.step_size);
            }
            ++steps;
            ++k;
            bool stop_here = ctx.transmittance < S(0.01);  // kMinTransmittance
            if (!forced) {
                if (stop_here) break;
                continue;
            }
            bool forced_stop = k >= forced_n;
            if (stop_here != forced_stop && !(forced_stop && cnt == 1))
                worst = std::fmax(worst, std::fabs(ctx.transmittance - S(0.01)) / S(0.01));
            if (forced_stop) break;
        }
    };
    num_steps = schedule(dist, 0);
    ctx.step_size = dist / S(num_steps);
    ctx.t = t1[0] + ctx.step_size * jitter;
    segment(num_steps, 0);
    S dist1 = t2[1] - t1[1];
    if (dist1 > S(0)) {
        dist1 = std::fmin(dist1, S(b.uMaxRaymarchDistance));
        uint32_t num_steps1 = schedule(dist1, 1);
        ctx.step_size = dist1 / S(num_steps1);
        ctx.t = t1[1] + ctx.step_size * jitter;
        segment(num_steps1, 1);
    }
    if (margin) *margin = worst;
    if (step_margin) *step_margin = worst_steps;
    CloudPixelOut<S> out;
    out.steps = steps;
    S average_t = ctx.weighted_t_sum == S(0) ? frag_dist : ctx.weighted_t_sum / ctx.transmittance_sum;
    out.distance = average_t;
    V3<S> average_pos = C.cam + view_dir * average_t;
    // GetSunVisibility(transmittance_texture, pos) (VolumetricCloudCommon.glsl:73-79)
    V3<S> up_dir = {average_pos.x, average_pos.y, average_pos.z + C.R};
    S rr = length(up_dir);
    up_dir = up_dir / rr;
    S mu_s = dot(C.sun, up_dir);
    V3<S> sun_vis = A.GetSunVisibility(im.transmittance, im.tw, im.th, rr, mu_s);
    // SunCosTheta (:85-88)
    S sun_cos = gclamp(dot(normalize(V3<S>{average_pos.x, average_pos.y, average_pos.z + C.R}), C.sun), S(0), S(1));
    V3<S> luminance = (ctx.sun_env[0] * S(b.uSunIlluminanceScale)) * sun_vis * A.solar_illuminance +
                      (ctx.sun_env[1] * std::pow(sun_cos, S(b.uEnvSunHeightCurveExp))) * cvt3<S>(b.uEnvColorScale);
    // GetAerialPerspective (VolumetricCloudCommon.glsl:81-98)
    S t_ap = average_t;
    if (r > A.top_radius) {
        S near_distance;
        if (FromSpaceIntersectTopAtmosphereBoundary(A, r, mu, near_distance))
            t_ap -= near_distance;
        else
            t_ap = 0;
    }
    S w = std::sqrt(t_ap / S(cb.uAerialPerspectiveLutMaxDistance));
    S uu = S(0.5) / S(im.aw) + u * (S(1) - S(1) / S(im.aw));
    S vv = S(0.5) / S(im.ah) + v * (S(1) - S(1) / S(im.ah));
    S ww = S(0.5) / S(im.ad) + w * (S(1) - S(1) / S(im.ad));
    V3<S> ap_t = tex3d_linear_clamp<S>(im.ap_transmittance, im.aw, im.ah, im.ad, uu, vv, ww);
    V3<S> ap_l = tex3d_linear_clamp<S>(im.ap_luminance, im.aw, im.ah, im.ad, uu, vv, ww);
    const FroxelBinding& froxel = BoundFroxel();
    if (froxel.data)  // VolumetricCloudRender.comp:198
        ap_l = ap_l * RayScatterVisibility<S>(froxel, u, v, average_t, S(cb.uInvShadowFroxelMaxDistance));
    luminance = luminance * ap_t + ap_l * (S(1) - ctx.transmittance);
    luminance = luminance * (S(1) / std::fmax(S(1e-5), S(1) - ctx.transmittance));
    S fade = smoothstep(S(b.uMaxVisibleDistance) * S(0.75), S(b.uMaxVisibleDistance), t1[0]);
    S T = gmix(ctx.transmittance, S(1), fade);
    luminance = luminance * (S(1) - T);
    out.rgba[0] = luminance.x;
    out.rgba[1] = luminance.y;
    out.rgba[2] = luminance.z;
    out.rgba[3] = T;
    return out;
}

}  // namespace oracle
