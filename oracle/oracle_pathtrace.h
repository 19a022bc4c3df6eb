// Test infrastructure (the oracle): CPU restatement of the voxel-cloud path
// tracer, shaders/SkyRendering/VolumetricCloudPathTracing.comp (Trace, main,
// display), with the materials' SampleSigmaT of oracle_cloud.h at the flat
// height of :89-91, the FP16 density grid of the path-tracing workload
// (level 0, LINEAR, CLAMP_TO_BORDER(0)) and its generator, and the
// environment term of the environment-luminance cube
// (AtmosphereRenderer.glsl:253-273) evaluated directly from the sky-view LUT.
// Tracking uses the reference's single majorant kSigmaTMax, so each sample
// consumes the reference's random sequence, or (when brick majorants are
// given) the library's B^3-brick majorant grid walked in texel space: the same
// tracking estimators with a piecewise-constant majorant, drawing random
// numbers in the same order as the library's brick tracker.  Only the tests
// and the benchmark's CPU baseline use this code.
#pragma once

#include "oracle_cloud.h"
#include "oracle_noise.h"

namespace oracle {

inline float half_to_float(uint16_t h) {
    uint32_t sign = uint32_t(h >> 15) << 31, exp = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    uint32_t f;
    if (exp == 0) {
        if (man == 0) {
            f = sign;
        } else {  // subnormal: renormalise
            int e = -1;
            do {
                man <<= 1;
                ++e;
            } while (!(man & 0x400u));
            f = sign | uint32_t(127 - 15 - e) << 23 | (man & 0x3ffu) << 13;
        }
    } else if (exp == 31) {
        f = sign | 0x7f800000u | man << 13;
    } else {
        f = sign | (exp + 127 - 15) << 23 | man << 13;
    }
    float out;
    std::memcpy(&out, &f, 4);
    return out;
}

// IEEE binary16 round-to-nearest-even of a finite float; overflow gives +-inf
// as the IEEE rule does.
inline uint16_t float_to_half_rne(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    uint32_t sign = (u >> 16) & 0x8000u;
    u &= 0x7fffffffu;
    if (u >= 0x477ff000u) return uint16_t(sign | 0x7c00u);  // rounds to >= 65536: inf
    if (u < 0x38800000u) {                                   // below the smallest normal half
        double r = std::nearbyint(double(std::fabs(x)) * 16777216.0);  // units of 2^-24, ties to even
        return uint16_t(sign | uint32_t(r));
    }
    uint32_t mant = u & 0x7fffffu, e = (u >> 23) - 127 + 15;
    uint32_t h = e << 10 | mant >> 13;
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // a carry moves into the exponent
    return uint16_t(sign | h);
}

// The path-tracing workload's density texel (the generator's formula, in FP32).
inline float VoxelGridDensity(int x, int y, int z, const int dim[3], const SkyhipVoxelGridCreateInfo& ci) {
    V3<float> p = {(float(x) + 0.5f) / float(dim[0]), (float(y) + 0.5f) / float(dim[1]),
                   (float(z) + 0.5f) / float(dim[2])};
    float res = 0.0f, a = 1.0f, sum_a = 0.0f;
    uint32_t f = ci.base_frequency;
    for (int o = 0; o < ci.octaves; ++o) {
        res += PerlinNoise<float>(p, f, ci.seed) * a;
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
    float falloff = std::fmax(1.0f - e, 0.0f);
    return std::fmax(std::exp(ci.lognormal_sigma * ci.fbm_inv_std * fbm) * falloff - ci.threshold, 0.0f);
}

// Default texels per brick edge of the library's majorant grid.
constexpr int kBrickTexels = 16;

// Brick majorant: max over the B^3 brick and a one-texel apron, clamped to the grid.
inline float BrickMajorant(const uint16_t* g, const int dim[3], int bx, int by, int bz, int B = kBrickTexels) {
    float m = 0.0f;
    int lo[3] = {bx * B - 1, by * B - 1, bz * B - 1};
    for (int z = std::max(lo[2], 0); z <= std::min(lo[2] + B + 1, dim[2] - 1); ++z)
        for (int y = std::max(lo[1], 0); y <= std::min(lo[1] + B + 1, dim[1] - 1); ++y)
            for (int x = std::max(lo[0], 0); x <= std::min(lo[0] + B + 1, dim[0] - 1); ++x)
                m = std::fmax(m, half_to_float(g[(size_t(z) * dim[1] + y) * dim[0] + x]));
    return m;
}

// The environment-luminance texel along world direction d (AtmosphereRenderer.glsl:253-273):
// the sky-view luminance, plus the lit ground times the sky-view transmittance where d hits
// the ground.
template <class S>
V3<S> EnvironmentLuminance(const Atmosphere<S>& A, const RenderBlock<S>& rb, const float* T, int tw, int th,
                           const float* sv_l, const float* sv_t, int sw, int sh, V3<S> d) {
    S cos_lat, cos_lon;
    GetCosLatLonFromViewDirection(rb, d, cos_lat, cos_lon);
    S r = rb.camera_earth_center_distance;
    S u, v;
    GetSkyViewTextureUvFromCosLatLon(A, sw, sh, r, cos_lat, cos_lon, u, v);
    V3<S> lum = tex2d_linear_clamp<S>(sv_l, sw, sh, u, v);
    V3<S> tr = tex2d_linear_clamp<S>(sv_t, sw, sh, u, v);
    if (A.RayIntersectsGround(r, cos_lat)) {
        S dist = A.DistanceToBottomAtmosphereBoundary(r, cos_lat);
        V3<S> ground = rb.camera_position + d * dist;
        lum += A.ComputeGroundLuminance(false, T, tw, th, rb.earth_center, ground, rb.sun_direction) * tr;
    }
    return lum;
}

// Trilinear CLAMP_TO_BORDER(0) sample of the FP16 grid at normalised (u, v, w).
template <class S>
S SampleGridF16(const uint16_t* g, const int dim[3], S u, S v, S w) {
    S c[3] = {u, v, w};
    int i0[3];
    S a[3];
    for (int k = 0; k < 3; ++k) {
        S x = c[k] * S(dim[k]) - S(0.5);
        S f = std::floor(x);
        a[k] = x - f;
        i0[k] = int(gclamp(f, S(-2), S(1 << 20)));
    }
    S r = 0;
    for (int corner = 0; corner < 8; ++corner) {
        int xi = i0[0] + (corner & 1), yi = i0[1] + ((corner >> 1) & 1), zi = i0[2] + ((corner >> 2) & 1);
        if (xi < 0 || yi < 0 || zi < 0 || xi >= dim[0] || yi >= dim[1] || zi >= dim[2]) continue;
        S wgt = ((corner & 1) ? a[0] : S(1) - a[0]) * ((corner & 2) ? a[1] : S(1) - a[1]) *
                ((corner & 4) ? a[2] : S(1) - a[2]);
        r += wgt * S(half_to_float(g[(size_t(zi) * dim[1] + yi) * dim[0] + xi]));
    }
    return r;
}

template <class S>
struct PathTracer {
    const Atmosphere<S>& A;
    const RenderBlock<S>& rb;
    const SkyhipCloudCommonBufferData& cb;
    const SkyhipCloudMaterial& mat;
    const SkyhipPathTracingParameters& pp;
    CloudCtx<S> cc;
    const float *T = nullptr, *sv_l = nullptr, *sv_t = nullptr, *ap_l = nullptr, *ap_t = nullptr;
    int tw = 0, th = 0, sw = 0, sh = 0, aw = 0, ah = 0, ad = 0;
    V3<S> cam, sun, box_min, box_max;
    S R, hb, ht;
    // Brick-majorant grid (VOXEL_F16 only): one majorant per brick^3 brick, in texel space
    // X = pos * tex_scale + tex_bias per axis (z measured from the cloud bottom).
    const float* majorants = nullptr;
    int brick = kBrickTexels;
    int nb[3] = {0, 0, 0};
    S tex_scale[3] = {0, 0, 0}, tex_bias[3] = {0, 0, 0}, grid_n[3] = {0, 0, 0};
    S majorant_scale = 0;
    static constexpr int kMaxTrackingSteps = 1 << 20;  // safety bounds, as the library's tracker
    static constexpr int kMaxSegments = 4096;

    PathTracer(const Atmosphere<S>& A_, const RenderBlock<S>& rb_, const SkyhipCloudCommonBufferData& cb_,
               const SkyhipCloudMaterial& m, const SkyhipPathTracingParameters& p)
        : A(A_), rb(rb_), cb(cb_), mat(m), pp(p), cc(&cb_, nullptr, &m) {
        cam = cvt3<S>(cb.uCameraPos);
        sun = cvt3<S>(cb.uSunDirection);
        R = cb.uEarthRadius;
        hb = cb.uBottomAltitude;
        ht = cb.uTopAltitude;
        // kCloudAABBMin/Max (VolumetricCloudPathTracing.comp:19-21)
        box_min = {-S(pp.region_box_half_width), -S(pp.region_box_half_width), hb};
        box_max = {S(pp.region_box_half_width), S(pp.region_box_half_width), ht};
    }

    void use_bricks(const float* m, int brick_size = kBrickTexels) {
        majorants = m;
        brick = brick_size;
        const SkyhipCloudVoxelBufferData& vb = mat.voxel;
        for (int k = 0; k < 3; ++k) {
            nb[k] = (mat.voxel_grid_dim[k] + brick - 1) / brick;
            grid_n[k] = S(mat.voxel_grid_dim[k]);
        }
        for (int k = 0; k < 2; ++k) {
            tex_scale[k] = S(vb.uSampleFrequency[k]) * grid_n[k];
            tex_bias[k] = S(vb.uSampleBias[k]) * grid_n[k];
        }
        tex_scale[2] = grid_n[2] / (ht - hb);
        tex_bias[2] = 0;
        majorant_scale = S(vb.uDensity);
    }

    // Majorant-segment walk over the brick grid.  The walk covers [-1, N + 1] texels in x
    // and y (the first and last bricks extend one texel beyond the grid, where
    // CLAMP_TO_BORDER samples are still non-zero); z is bounded by the cloud slab.
    struct Walk {
        S org[3], k[3], inv_k[3], tn[3], t_end;
        int b[3];
    };
    S next_boundary(const Walk& w, int i) const {
        if (w.k[i] > S(0)) return w.b[i] < nb[i] - 1 ? (S(brick * (w.b[i] + 1)) - w.org[i]) * w.inv_k[i] : S(3.0e38);
        if (w.k[i] < S(0)) return w.b[i] > 0 ? (S(brick * w.b[i]) - w.org[i]) * w.inv_k[i] : S(3.0e38);
        return S(3.0e38);
    }
    bool walk_begin(Walk& w, V3<S> o, V3<S> d, S& t, S t_stop) const {
        const S oo[3] = {o.x, o.y, o.z - hb}, dd[3] = {d.x, d.y, d.z};
        S lo = t, hi = t_stop;
        for (int i = 0; i < 3; ++i) {
            w.org[i] = oo[i] * tex_scale[i] + tex_bias[i];
            w.k[i] = dd[i] * tex_scale[i];
            if (std::fabs(w.k[i]) < S(1e-30)) w.k[i] = 0;
            w.inv_k[i] = w.k[i] != S(0) ? S(1) / w.k[i] : S(0);
            if (i < 2) {
                S e0 = (S(-1) - w.org[i]) * w.inv_k[i], e1 = (grid_n[i] + S(1) - w.org[i]) * w.inv_k[i];
                if (w.k[i] == S(0)) {
                    if (w.org[i] < S(-1) || w.org[i] > grid_n[i] + S(1)) return false;
                } else {
                    lo = std::fmax(lo, std::fmin(e0, e1));
                    hi = std::fmin(hi, std::fmax(e0, e1));
                }
            }
        }
        if (!(lo < hi)) return false;
        t = lo;
        w.t_end = hi;
        for (int i = 0; i < 3; ++i) {
            S x = t * w.k[i] + w.org[i];
            w.b[i] = std::min(std::max(int(std::floor(x / S(brick))), 0), nb[i] - 1);
            w.tn[i] = next_boundary(w, i);
        }
        return true;
    }
    S segment_end(const Walk& w) const { return std::fmin(std::fmin(w.tn[0], w.tn[1]), std::fmin(w.tn[2], w.t_end)); }
    S segment_majorant(const Walk& w) const {
        return S(majorants[(size_t(w.b[2]) * nb[1] + w.b[1]) * nb[0] + w.b[0]]) * majorant_scale;
    }
    void advance(Walk& w) const {
        int i = (w.tn[0] <= w.tn[1] && w.tn[0] <= w.tn[2]) ? 0 : (w.tn[1] <= w.tn[2] ? 1 : 2);
        w.b[i] = std::min(std::max(w.b[i] + (w.k[i] > S(0) ? 1 : -1), 0), nb[i] - 1);
        w.tn[i] = next_boundary(w, i);
    }

    // Delta tracking over [t0, t1] with the brick majorants; true with t at a real collision.
    bool delta_track_bricks(uint32_t& seed, V3<S> o, V3<S> d, S t0, S t1, S& t) const {
        Walk w;
        t = t0;
        if (!walk_begin(w, o, d, t, t1)) return false;
        int it = 0;
        for (int seg = 0; seg < kMaxSegments; ++seg) {
            S seg_end = segment_end(w), mu = segment_majorant(w);
            if (mu > S(0)) {
                for (; it < kMaxTrackingSteps; ++it) {
                    S tt = t + -std::log(S(1) - random01(seed)) / mu;
                    if (!(tt < seg_end)) break;
                    t = tt;
                    if (random01(seed) < sigma(o + d * t) / mu) return true;
                }
                if (it >= kMaxTrackingSteps) return false;
            }
            if (!(seg_end < w.t_end)) return false;
            t = seg_end;
            advance(w);
        }
        return false;
    }

    // Ratio tracking with the brick majorants.
    S ratio_track_bricks(uint32_t seed, V3<S> o, V3<S> d, S t0, S t1) const {
        S tr = 1;
        Walk w;
        S t = t0;
        if (!walk_begin(w, o, d, t, t1)) return tr;
        int it = 0;
        for (int seg = 0; seg < kMaxSegments; ++seg) {
            S seg_end = segment_end(w), mu = segment_majorant(w);
            if (mu > S(0)) {
                for (; it < kMaxTrackingSteps; ++it) {
                    S tt = t + -std::log(S(1) - random01(seed)) / mu;
                    if (!(tt < seg_end)) break;
                    t = tt;
                    tr *= S(1) - std::fmax(S(0), sigma(o + d * t) / mu);
                }
                if (it >= kMaxTrackingSteps || tr == S(0)) break;
            }
            if (!(seg_end < w.t_end)) break;
            t = seg_end;
            advance(w);
        }
        return gclamp(tr, S(0), S(1));
    }

    uint32_t prng(uint32_t x) const { return pp.prng == SKYHIP_PRNG_WANG_HASH ? WangHash(x) : PCGHash(x); }
    // Random01 (:44-48): returns the current state, then advances it.
    S random01(uint32_t& seed) const {
        S r = S(seed) / S(4294967296.0);
        seed = prng(seed);
        return r;
    }

    // CloudRegionIntersect (:57-75)
    void region(V3<S> o, V3<S> d, S& t0, S& t1) const {
        t0 = 0, t1 = S(1e7);
        for (int i = 0; i < 3; ++i) {
            S a = (box_min[i] - o[i]) / d[i], b = (box_max[i] - o[i]) / d[i];
            t0 = std::fmax(t0, std::fmin(a, b));
            t1 = std::fmin(t1, std::fmax(a, b));
        }
    }

    // SampleSigmaT(P) with the flat height (:89-91).
    S sigma(V3<S> P) const {
        S h01 = gclamp((P.z - hb) / (ht - hb), S(0), S(1));
        if (mat.kind == SKYHIP_MATERIAL_VOXEL_F16) {
            const SkyhipCloudVoxelBufferData& p = mat.voxel;
            S u = P.x * S(p.uSampleFrequency[0]) + S(p.uSampleBias[0]);
            S v = P.y * S(p.uSampleFrequency[1]) + S(p.uSampleBias[1]);
            return SampleGridF16<S>(mat.voxel_grid, mat.voxel_grid_dim, u, v, h01) * S(p.uDensity);
        }
        return cc.SampleSigmaT(P, h01);
    }

    // GetPhase (:93-96)
    S phase(S c) const {
        return gmix(CloudCtx<S>::HenyeyGreenstein(c, S(pp.back_phase_g)),
                    CloudCtx<S>::HenyeyGreenstein(c, S(pp.forward_phase_g)), S(pp.forward_scattering_ratio));
    }

    // CreateOrthonormalBasis, shaders/Base/Common.glsl:44-51 (Duff et al. 2017)
    static void onb(V3<S> N, V3<S>& t0, V3<S>& t1) {
        S s = N.z >= S(0) ? S(1) : S(-1);
        S a = S(-1) / (s + N.z);
        S b = N.x * N.y * a;
        t0 = {S(1) + s * N.x * N.x * a, s * b, -s * N.x};
        t1 = {b, s + N.y * N.y * a, -N.y};
    }

    // HenyeyGreensteinInvertcdf, VolumetricCloudCommon.glsl:65-71
    static S hg_invert_cdf(S xi, S g) {
        S one_plus_g2 = S(1) + g * g, one_minus_g2 = S(1) - g * g, one_over_2g = S(0.5) / g;
        S t = one_minus_g2 / (S(1) - g + S(2) * g * xi);
        return one_over_2g * (one_plus_g2 - t * t);
    }

    // GenerateHGSample (:98-117)
    V3<S> hg_sample(uint32_t& seed, V3<S> d, S& value_over_pdf) const {
        if (pp.importance_sampling) {
            S g = random01(seed) < S(pp.forward_scattering_ratio) ? S(pp.forward_phase_g) : S(pp.back_phase_g);
            S cos_theta = hg_invert_cdf(random01(seed), g);
            S sin_theta = std::sqrt(gclamp(S(1) - cos_theta * cos_theta, S(0), S(1)));
            V3<S> t0, t1;
            onb(d, t0, t1);
            S ph = S(2) * S(kPi) * random01(seed);
            value_over_pdf = S(1);
            return sin_theta * std::sin(ph) * t0 + sin_theta * std::cos(ph) * t1 + cos_theta * d;
        }
        // UniformSphereSample (:50-55)
        S ph = S(2) * S(kPi) * random01(seed);
        S cos_theta = S(1) - S(2) * random01(seed);
        S sin_theta = std::sqrt(gclamp(S(1) - cos_theta * cos_theta, S(0), S(1)));
        V3<S> dir = {std::cos(ph) * sin_theta, std::sin(ph) * sin_theta, cos_theta};
        value_over_pdf = phase(dot(d, dir)) / (S(1) / (S(4) * S(kPi)));
        return dir;
    }

    // GenerateLambertSample (:119-129)
    V3<S> lambert_sample(uint32_t& seed, V3<S> N) const {
        S sin_theta = std::sqrt(random01(seed));
        S cos_theta = std::sqrt(gclamp(S(1) - sin_theta * sin_theta, S(0), S(1)));
        V3<S> t0, t1;
        onb(N, t0, t1);
        S ph = S(2) * S(kPi) * random01(seed);
        return sin_theta * std::sin(ph) * t0 + sin_theta * std::cos(ph) * t1 + cos_theta * N;
    }

    // GetSunIlluminance (:131-133) with GetSunVisibility(pos), VolumetricCloudCommon.glsl:73-79.
    V3<S> sun_illuminance(V3<S> pos) const {
        V3<S> up = {pos.x, pos.y, pos.z + R};
        S r = length(up);
        up = up / r;
        return A.GetSunVisibility(T, tw, th, r, dot(sun, up)) * A.solar_illuminance;
    }

    // TransmittanceEstimation (:135-151): ratio tracking on a copy of the RNG state.
    S transmittance_estimation(uint32_t seed, V3<S> o, V3<S> d) const {
        S tr = 1;
        S t0, t1;
        region(o, d, t0, t1);
        if (t0 >= t1) return tr;
        if (majorants) return ratio_track_bricks(seed, o, d, t0, t1);
        S smax = S(pp.sigma_t_max);
        S t = t0;
        for (;;) {
            t += -std::log(S(1) - random01(seed)) / smax;
            if (t > t1) break;
            tr *= S(1) - std::fmax(S(0), sigma(o + d * t) / smax);
        }
        return gclamp(tr, S(0), S(1));
    }

    // Delta tracking over [t0, t1] (:182-196): true with t at a real collision.
    bool delta_track(uint32_t& seed, V3<S> o, V3<S> d, S t0, S t1, S& t) const {
        if (majorants) return delta_track_bricks(seed, o, d, t0, t1, t);
        const S smax = S(pp.sigma_t_max);
        t = t0;
        if (smax <= S(0)) return false;
        for (;;) {
            t += -std::log(S(1) - random01(seed)) / smax;
            if (t > t1) return false;
            S s = sigma(o + d * t);
            if (random01(seed) < s / smax) return true;
        }
    }

    // SampleLuminanceFromLight (:153-158); the estimator gets the state by value.
    V3<S> light(uint32_t seed, V3<S> pos, V3<S> bsdf) const {
        return transmittance_estimation(seed, pos, sun) * sun_illuminance(pos) * bsdf;
    }

    // Environment-luminance texel for a local-frame direction: AtmosphereRenderer.glsl:253-273
    // at direction kModelMatrix3 * d.
    V3<S> environment(V3<S> local_dir) const {
        const float* m = pp.model3;
        V3<S> d = {S(m[0]) * local_dir.x + S(m[3]) * local_dir.y + S(m[6]) * local_dir.z,
                   S(m[1]) * local_dir.x + S(m[4]) * local_dir.y + S(m[7]) * local_dir.z,
                   S(m[2]) * local_dir.x + S(m[5]) * local_dir.y + S(m[8]) * local_dir.z};
        return EnvironmentLuminance(A, rb, T, tw, th, sv_l, sv_t, sw, sh, d);
    }

    // Trace (:160-249); returns L and writes has_scattered / scattered_t.
    V3<S> trace(uint32_t& seed, V3<S> view, bool& has_scattered, S& scattered_t) const {
        V3<S> L = v3(S(0)), throughput = v3(S(1));
        has_scattered = false;
        scattered_t = 0;
        V3<S> o = cam, d = view;
        S c0, c1;
        region(o, d, c0, c1);
        if (c0 >= c1) return L;
        o = o + c0 * d;
        for (int istep = 0;
             istep < pp.max_bounces && std::fmax(throughput.x, std::fmax(throughput.y, throughput.z)) > S(0);
             ++istep) {
            S t0, t1;
            region(o, d, t0, t1);
            if (t0 >= t1) break;
            S t;
            if (!delta_track(seed, o, d, t0, t1, t)) {
                if (pp.environment_lighting == SKYHIP_ENV_OFF) break;
                if (!has_scattered) break;
                if (pp.environment_lighting == SKYHIP_ENV_CONST_ENVIRONMENT_MAP) {
                    L += throughput * environment(d);
                    break;
                }
                V3<S> up = {o.x, o.y, o.z + R};
                S r = length(up);
                up = up / r;
                S mu = dot(d, up);
                if (!A.RayIntersectsGround(r, mu)) {
                    L += throughput * environment(d);
                    break;
                }
                o = o + d * A.DistanceToBottomAtmosphereBoundary(r, mu);
                V3<S> N = normalize(V3<S>{o.x, o.y, o.z + R});
                V3<S> bsdf = S(1.0 / kPi) * A.ground_albedo;
                L += throughput * light(seed, o, bsdf * dot(N, sun));
                if (pp.environment_lighting == SKYHIP_ENV_GROUND_SINGLE_BOUNCE) break;
                d = lambert_sample(seed, N);
                throughput *= A.ground_albedo;
            } else {
                if (!has_scattered) scattered_t = distance(cam, o);  // the segment start, as :232-233
                has_scattered = true;
                o = o + d * t;
                // A real collision scatters with probability albedo (the reference's albedo is 1):
                // both the sun estimate at this vertex and the continuation carry it.
                throughput *= S(pp.albedo);
                L += throughput * light(seed, o, v3(phase(dot(d, sun))));
                S vop;
                d = hg_sample(seed, d, vop);
                throughput *= vop;
            }
        }
        return L;
    }

    // main (:255-284): one sample of pixel (x, y) of a width x height image at frame `frame`.
    void sample(int x, int y, int width, int height, uint32_t frame, S out[4]) const {
        uint32_t seed = prng(prng(prng(uint32_t(x)) + uint32_t(y)) + frame);
        S u = (S(x) + S(0.5)) / S(width), v = (S(y) + S(0.5)) / S(height);
        V3<S> fp = projective_mul<S>(cb.uInvMVP, {u * S(2) - S(1), v * S(2) - S(1), S(1)});
        V3<S> view = normalize(fp - cam);
        bool hs;
        S st;
        V3<S> L = trace(seed, view, hs, st);
        if (hs) {  // GetAerialPerspective (VolumetricCloudCommon.glsl:81-98) and the froxel (:277)
            S r = cam.z + R, mu = view.z;
            S t = st;
            if (r > A.top_radius) {
                S nd;
                if (FromSpaceIntersectTopAtmosphereBoundary(A, r, mu, nd))
                    t -= nd;
                else
                    t = 0;
            }
            S w = std::sqrt(t / S(cb.uAerialPerspectiveLutMaxDistance));
            S uu = S(0.5) / S(aw) + u * (S(1) - S(1) / S(aw)), vv = S(0.5) / S(ah) + v * (S(1) - S(1) / S(ah));
            S ww = S(0.5) / S(ad) + w * (S(1) - S(1) / S(ad));
            V3<S> apt = tex3d_linear_clamp<S>(ap_t, aw, ah, ad, uu, vv, ww);
            V3<S> apl = tex3d_linear_clamp<S>(ap_l, aw, ah, ad, uu, vv, ww);
            const FroxelBinding& froxel = BoundFroxel();
            if (froxel.data) apl = apl * RayScatterVisibility<S>(froxel, u, v, st, S(cb.uInvShadowFroxelMaxDistance));
            L = L * apt + apl;
        }
        out[0] = L.x, out[1] = L.y, out[2] = L.z, out[3] = hs ? S(0) : S(1);
    }
};

}  // namespace oracle
