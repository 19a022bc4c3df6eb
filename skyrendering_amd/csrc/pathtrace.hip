// Voxel-cloud path tracer for gfx950 (VolumetricCloudPathTracing.comp):
// delta tracking for free paths, ratio tracking for the sun's shadow rays,
// Henyey-Greenstein mixture and Lambert sampling, ground bounces, the
// environment term from the sky-view LUT and aerial perspective at the first
// scattering event.
//
// One path per lane; a wave64, one workgroup, covers an 8x8 pixel tile and
// loops over the requested frames, adding each frame's
// sample to the pixel's accumulator in frame order as the reference's
// per-frame imageLoad/imageStore does.  Tracking runs either with the
// reference's single majorant kSigmaTMax, drawing the reference's random
// sequence, or over the B^3-brick majorant grid of an FP16 density volume (B = 16
// by default),
// walked with a 3D DDA in texel space: empty bricks cost one majorant fetch,
// and inside occupied bricks the tentative-collision rate follows the local
// maximum instead of the global one.
#include "cloud_device.h"
#include "sky_env.h"

namespace skyhip {

int build_material(const SkyhipCloudMaterial* m, MaterialDev& d);
int validate_atmosphere(const SkyhipAtmosphereBufferData* b);
int validate_lut(const void* p, int w, int h, const char* what);
int validate_render(const SkyhipAtmosphereRenderBufferData* b);

// Safety bound on tentative collisions per tracking call (never reached by a
// finite majorant over a bounded region; keeps every wave's loop finite).
constexpr int kMaxTrackingSteps = 1 << 20;
// Safety bound on majorant segments per walk.  A walk crosses at most the sum of the
// brick-grid dimensions plus two: grids of up to 4096 texels a side in bricks of at least 4
// texels give at most 3 * 1024 + 2 segments.
constexpr int kMaxSegments = 4096;

struct PtArgs {
    MaterialDev mat;
    AtmDev A;
    // Cloud local frame (VolumetricCloudCommon.glsl block)
    F3 cam, sun;
    float R, hb, ht, inv_thickness, half_width;
    float inv_mvp[16];
    float ap_max_distance;
    // PathTracing::InitParam
    int max_bounces, importance_sampling, prng, env;
    float gf, gb, ratio, sigma_t_max, inv_sigma_t_max, albedo;
    F3 model0, model1, model2;  // kModelMatrix3 columns
    // Atmosphere renderer frame for the environment term
    EnvFrame sky;
    // LUTs
    const float4* T;
    int tw, th;
    const float4* sv_l;
    const float4* sv_t;
    int sw, sh;
    const float4* ap_l;
    const float4* ap_t;
    int aw, ah, ad;
    // Brick majorants: texel-space X = pos * tex_scale + tex_bias per axis (z on the flat height).
    // The walks read inv_majorants[b] = 1 / (majorants[b] * majorant_scale), 0 for an empty brick
    // (inv_majorant_kernel, before every launch).
    const float* majorants;
    const float* inv_majorants;
    int nb[3];
    int brick;        // brick edge B in texels (a power of two)
    float inv_brick;  // 1 / B, exact
    float tex_scale[3], tex_bias[3], grid_n[3];
    float majorant_scale;
    // The grid's cell-corner copy (one 16-byte load per density lookup), or null
    const uint4* cells;
    // The context's bound ray-scatter-visibility froxel, or null (visibility 1)
    const uint16_t* froxel;
    int fw, fh, fd;
    float inv_froxel_max_distance;
    // Output
    int region[4];
    uint32_t first_frame;
    int frames;
    int width, height;
    float4* accum;
    uint8_t* mask;
    unsigned long long* counters;  // samples, density lookups, majorant segments, path vertices, shadow rays
    unsigned* queue;               // path_trace_refill_kernel's tile counter (zero at launch)
};

struct PtCount {
    uint32_t lookups = 0, segments = 0, vertices = 0, shadow_rays = 0;
};

__device__ __forceinline__ uint32_t wang_hash_pt(uint32_t seed) {  // Noise.glsl:1-9
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed;
}
__device__ __forceinline__ uint32_t pcg_hash(uint32_t seed) {  // Noise.glsl:11-15
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Random01 (VolumetricCloudPathTracing.comp:44-48): the current state as a float in [0, 1], then
// the state hashed forward.
struct Rng {
    uint32_t s;
    int pcg;
    __device__ __forceinline__ uint32_t hash(uint32_t x) const { return pcg ? pcg_hash(x) : wang_hash_pt(x); }
    __device__ __forceinline__ float next() {
        float r = float(s) * 2.3283064365386963e-10f;
        s = hash(s);
        return r;
    }
};

// InfiniteTransmittanceIS numerator: -log(1 - xi).
// -ln(1 - xi) for Random01's xi = seed * 2^-32.  The float conversion rounds seeds from
// 2^32 - 128 up to xi = 1, where the shader's log(1 - xi) is -inf and the step +inf (it leaves
// the segment); that case is returned as +inf explicitly, because the product below would turn
// -inf into NaN (a NaN step never leaves the single-majorant loops before their safety cap).
// Otherwise 1 - xi is a normal float in [2^-32, 1], so the hardware log2 (v_log_f32) needs none
// of the library logf's denormal and infinity handling.  ln = log2 * ln2, with ln2 split into a
// float kLn2Hi just below it and the remainder kLn2Lo, and the product's rounding error
// recovered with an FMA, so the result keeps FP32 precision.
__device__ __forceinline__ float exp_sample(float xi) {
    constexpr float kLn2Hi = 0x1.62e42ep-1f;     // 0.693147122...
    constexpr float kLn2Lo = 0x1.efa39ep-25f;    // ln 2 - kLn2Hi = 5.77e-8
    const float x = 1.0f - xi;
    const float l2 = __builtin_amdgcn_logf(fmaxf(x, 0x1p-32f));
    const float hi = l2 * kLn2Hi;
    const float r = -(hi + (fmaf(l2, kLn2Hi, -hi) + l2 * kLn2Lo));
    return x > 0.0f ? r : __builtin_inff();
}

// CloudRegionIntersect (:57-75) on the slab +-half_width x [hb, ht].
__device__ __forceinline__ void region(const PtArgs& a, F3 o, F3 d, float& t0, float& t1) {
    float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
    float x0 = (-a.half_width - o.x) * ix, x1 = (a.half_width - o.x) * ix;
    float y0 = (-a.half_width - o.y) * iy, y1 = (a.half_width - o.y) * iy;
    float z0 = (a.hb - o.z) * iz, z1 = (a.ht - o.z) * iz;
    t0 = fmaxf(fmaxf(0.0f, fminf(x0, x1)), fmaxf(fminf(y0, y1), fminf(z0, z1)));
    t1 = fminf(fminf(1e7f, fmaxf(x0, x1)), fminf(fmaxf(y0, y1), fmaxf(z0, z1)));
}

template <int kMat>
__device__ __forceinline__ float sigma_at(const PtArgs& a, F3 p) {
    // SampleSigmaT(P) with the flat height01 of :89-91
    return sample_sigma_t<kMat>(a.mat, p, clamp01((p.z - a.hb) * a.inv_thickness), a.cam);
}

// Default texels per brick edge of the majorant grid.  Measured on the 1024^2 workload (512^3
// grid): 8 -> 11.25 ms per frame (82 M lookups, 240 M segments), 12 -> 10.76 ms (102 M, 164 M),
// 16 -> 10.70-10.86 ms (119 M, 123 M), 20 -> 12.02 ms, 24 -> 12.84 ms, 32 -> 14.68 ms: past 16
// the looser majorants cost more null collisions than the shorter walk saves.
constexpr int kDefaultBrick = 16;

// Majorant-segment walker over the brick grid: the ray's texel-space position is
// X(t) = org + t k per axis.  Bricks are a.brick texels wide; the first and last brick of
// the x and y axes extend one texel beyond the grid, where CLAMP_TO_BORDER samples
// are still non-zero, so the walk covers [-1, N + 1] there and [0, N] in z.
struct BrickWalk {
    float org[3], k[3], inv_k[3];
    int b[3];
    float tn[3];
    float t_end;

    // Sets up the walk over [t, t_end] clipped to the grid's support; false if empty.
    __device__ bool begin(const PtArgs& a, F3 o, F3 d, float& t, float t_stop) {
        const float oo[3] = {o.x, o.y, o.z - a.hb}, dd[3] = {d.x, d.y, d.z};
        float lo = t, hi = t_stop;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            org[i] = oo[i] * a.tex_scale[i] + a.tex_bias[i];
            k[i] = dd[i] * a.tex_scale[i];
            if (fabsf(k[i]) < 1e-30f) k[i] = 0.0f;  // keeps every boundary time finite
            inv_k[i] = k[i] != 0.0f ? 1.0f / k[i] : 0.0f;
            if (i < 2) {  // clip to [-1, N + 1]
                float e0 = (-1.0f - org[i]) * inv_k[i], e1 = (a.grid_n[i] + 1.0f - org[i]) * inv_k[i];
                if (k[i] == 0.0f) {
                    if (org[i] < -1.0f || org[i] > a.grid_n[i] + 1.0f) return false;
                } else {
                    lo = fmaxf(lo, fminf(e0, e1));
                    hi = fminf(hi, fmaxf(e0, e1));
                }
            }
        }
        if (!(lo < hi)) return false;
        t = lo;
        t_end = hi;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float x = fmaf(t, k[i], org[i]);
            b[i] = min(max(int(floorf(x * a.inv_brick)), 0), a.nb[i] - 1);
            tn[i] = next_boundary(a, i);
        }
        return true;
    }
    __device__ __forceinline__ float next_boundary(const PtArgs& a, int i) const {
        if (k[i] > 0.0f) return b[i] < a.nb[i] - 1 ? (float(a.brick * (b[i] + 1)) - org[i]) * inv_k[i] : 3.0e38f;
        if (k[i] < 0.0f) return b[i] > 0 ? (float(a.brick * b[i]) - org[i]) * inv_k[i] : 3.0e38f;
        return 3.0e38f;
    }
    __device__ __forceinline__ float segment_end() const { return fminf(fminf(tn[0], tn[1]), fminf(tn[2], t_end)); }
    __device__ __forceinline__ int brick(const PtArgs& a) const { return (b[2] * a.nb[1] + b[1]) * a.nb[0] + b[0]; }
    template <int i>
    __device__ __forceinline__ void step(const PtArgs& a) {
        b[i] = min(max(b[i] + (k[i] > 0.0f ? 1 : -1), 0), a.nb[i] - 1);
        tn[i] = next_boundary(a, i);
    }
    // Moves into the next brick along the axis whose boundary comes first (constant
    // indices only, so the walker stays in registers).
    __device__ __forceinline__ void advance(const PtArgs& a) {
        if (tn[0] <= tn[1] && tn[0] <= tn[2])
            step<0>(a);
        else if (tn[1] <= tn[2])
            step<1>(a);
        else
            step<2>(a);
    }
};

// The voxel material's density at ray parameter t of a brick walk, sampled in the walk's texel
// space: the walk's X(t) = org + t k is u N per axis (z from the layer bottom), so the
// trilinear texel coordinate is X(t) - 0.5; equal to sigma_at(o + d t) in exact arithmetic.
// kCells reads the grid's cell-corner copy (same texels, same arithmetic).
template <bool kCells>
__device__ __forceinline__ float sigma_walk(const PtArgs& a, const BrickWalk& w, float t) {
    const float x = fmaf(t, w.k[0], w.org[0]) - 0.5f, y = fmaf(t, w.k[1], w.org[1]) - 0.5f,
                z = fmaf(t, w.k[2], w.org[2]) - 0.5f;
    const float g = kCells ? grid_cells_trilinear_texel(a.cells, a.mat.grid.dim, x, y, z)
                           : grid_f16_trilinear_texel(a.mat.grid, x, y, z);
    return g * a.mat.voxel.uDensity;
}

// Delta tracking over [t0, t1] (:182-196); returns true with t at a real collision.  kFlat
// selects the loop form of the brick walk (below); both make the same draws.
template <int kMat, bool kBricks, bool kCount, bool kCells, bool kFlat>
__device__ bool delta_track(const PtArgs& a, Rng& rng, F3 o, F3 d, float t0, float t1, float& t, PtCount& wc) {
    if (!kBricks) {
        if (a.sigma_t_max <= 0.0f) return false;
        t = t0;
        for (int it = 0; it < kMaxTrackingSteps; ++it) {
            t += exp_sample(rng.next()) * a.inv_sigma_t_max;
            if (t > t1) return false;
            float s = sigma_at<kMat>(a, o + d * t);
            if (kCount) wc.lookups++;
            if (rng.next() < s * a.inv_sigma_t_max) return true;
        }
        return false;
    }
    BrickWalk w;
    t = t0;
    if (!w.begin(a, o, d, t, t1)) return false;
    if constexpr (!kFlat) {
        // A loop over bricks around a loop over the tentative collisions in each.
        int it = 0;
        for (int seg = 0; seg < kMaxSegments; ++seg) {
            const float seg_end = w.segment_end();
            const float inv_mu = a.inv_majorants[w.brick(a)];
            if (kCount) wc.segments++;
            if (inv_mu > 0.0f) {
                for (; it < kMaxTrackingSteps; ++it) {
                    float tt = t + exp_sample(rng.next()) * inv_mu;
                    if (!(tt < seg_end)) break;
                    t = tt;
                    float s = sigma_walk<kCells>(a, w, t);
                    if (kCount) wc.lookups++;
                    if (rng.next() < s * inv_mu) return true;
                }
                if (it >= kMaxTrackingSteps) return false;
            }
            if (!(seg_end < w.t_end)) return false;
            t = seg_end;
            w.advance(a);
        }
        return false;
    }
    // One loop over tentative collisions and brick crossings together: an iteration draws the next
    // tentative collision in the current brick and looks it up, or, when the draw falls past the
    // brick, moves to the next brick.  Lanes of a wave then differ only in how many iterations
    // they take, where the nested form makes the wave wait in every brick for its lane with the
    // most collisions there, but every iteration carries both bodies.  Same draws in the same
    // order as the nested form (a draw past the brick is discarded), so the same samples.
    float seg_end = w.segment_end();
    float inv_mu = a.inv_majorants[w.brick(a)];
    if (kCount) wc.segments++;
    int it = 0, seg = 0;
    for (;;) {
        bool cross = true;
        if (inv_mu > 0.0f && it < kMaxTrackingSteps) {
            const float tt = t + exp_sample(rng.next()) * inv_mu;
            if (tt < seg_end) {
                cross = false;
                t = tt;
                ++it;
                const float s = sigma_walk<kCells>(a, w, t);
                if (kCount) wc.lookups++;
                if (rng.next() < s * inv_mu) return true;
            }
        }
        if (cross) {
            if (inv_mu > 0.0f && it >= kMaxTrackingSteps) return false;
            if (!(seg_end < w.t_end) || ++seg >= kMaxSegments) return false;
            t = seg_end;
            w.advance(a);
            seg_end = w.segment_end();
            inv_mu = a.inv_majorants[w.brick(a)];
            if (kCount) wc.segments++;
        }
    }
}

// TransmittanceEstimation (:135-151): ratio tracking toward the sun on a copy of the state
// (kFlat as in delta_track).
template <int kMat, bool kBricks, bool kCount, bool kCells, bool kFlat>
__device__ float ratio_track(const PtArgs& a, Rng rng, F3 o, F3 d, PtCount& wc) {
    float t0, t1;
    region(a, o, d, t0, t1);
    float tr = 1.0f;
    if (t0 >= t1) return tr;
    if (kCount) wc.shadow_rays++;
    if (!kBricks) {
        float t = t0;
        for (int it = 0; it < kMaxTrackingSteps; ++it) {
            t += exp_sample(rng.next()) * a.inv_sigma_t_max;
            if (t > t1) break;
            float s = sigma_at<kMat>(a, o + d * t);
            if (kCount) wc.lookups++;
            tr *= 1.0f - fmaxf(0.0f, s * a.inv_sigma_t_max);
        }
        return clamp01(tr);
    }
    // Two tentative collisions per lookup iteration, their lookups issued together: the second is
    // drawn ahead, and when it falls past the brick (or the step bound) the state goes back to
    // before its draw, so the next iteration draws it again, overshoots and crosses as the
    // one-at-a-time loop does.  Same draws, same products, same order, in either loop form.
    BrickWalk w;
    float t = t0;
    if (!w.begin(a, o, d, t, t1)) return tr;
    if constexpr (!kFlat) {
        int it = 0;
        for (int seg = 0; seg < kMaxSegments; ++seg) {
            const float seg_end = w.segment_end();
            const float inv_mu = a.inv_majorants[w.brick(a)];
            if (kCount) wc.segments++;
            if (inv_mu > 0.0f) {
                while (it < kMaxTrackingSteps) {
                    const float tt = t + exp_sample(rng.next()) * inv_mu;
                    if (!(tt < seg_end)) break;
                    const uint32_t before_second = rng.s;
                    const float tt2 = tt + exp_sample(rng.next()) * inv_mu;
                    const bool two = tt2 < seg_end && it + 1 < kMaxTrackingSteps;
                    const float s = sigma_walk<kCells>(a, w, tt);
                    const float s2 = sigma_walk<kCells>(a, w, two ? tt2 : tt);
                    tr *= 1.0f - fmaxf(0.0f, s * inv_mu);
                    if (two) {
                        tr *= 1.0f - fmaxf(0.0f, s2 * inv_mu);
                        t = tt2;
                        it += 2;
                    } else {
                        rng.s = before_second;
                        t = tt;
                        it += 1;
                    }
                    if (kCount) wc.lookups += two ? 2u : 1u;
                }
                if (it >= kMaxTrackingSteps || tr == 0.0f) break;
            }
            if (!(seg_end < w.t_end)) break;
            t = seg_end;
            w.advance(a);
        }
        return clamp01(tr);
    }
    float seg_end = w.segment_end();
    float inv_mu = a.inv_majorants[w.brick(a)];
    if (kCount) wc.segments++;
    int it = 0, seg = 0;
    for (;;) {
        bool cross = true;
        if (inv_mu > 0.0f && it < kMaxTrackingSteps) {
            const float tt = t + exp_sample(rng.next()) * inv_mu;
            if (tt < seg_end) {
                cross = false;
                const uint32_t before_second = rng.s;
                const float tt2 = tt + exp_sample(rng.next()) * inv_mu;
                const bool two = tt2 < seg_end && it + 1 < kMaxTrackingSteps;
                const float s = sigma_walk<kCells>(a, w, tt);
                const float s2 = sigma_walk<kCells>(a, w, two ? tt2 : tt);
                tr *= 1.0f - fmaxf(0.0f, s * inv_mu);
                if (two) {
                    tr *= 1.0f - fmaxf(0.0f, s2 * inv_mu);
                    t = tt2;
                    it += 2;
                } else {
                    rng.s = before_second;
                    t = tt;
                    it += 1;
                }
                if (kCount) wc.lookups += two ? 2u : 1u;
            }
        }
        if (cross) {
            if (inv_mu > 0.0f && (it >= kMaxTrackingSteps || tr == 0.0f)) break;
            if (!(seg_end < w.t_end) || ++seg >= kMaxSegments) break;
            t = seg_end;
            w.advance(a);
            seg_end = w.segment_end();
            inv_mu = a.inv_majorants[w.brick(a)];
            if (kCount) wc.segments++;
        }
    }
    return clamp01(tr);
}

// GetSunIlluminance (:131-133) at a local-frame point.
__device__ inline F3 sun_illuminance(const PtArgs& a, F3 p) {
    // r^2 - R_b^2 with the cloud's planet radius R: x^2 + y^2 + z (z + 2R) + (R - R_b)(R + R_b)
    float q = p.x * p.x + p.y * p.y + p.z * (p.z + 2.0f * a.R) + (a.R - a.A.rb) * (a.R + a.A.rb);
    float r = sqrtf(fmaxf(q + a.A.rb * a.A.rb, 0.0f));
    float alt = q / (r + a.A.rb);
    float mu_s = (a.sun.x * p.x + a.sun.y * p.y + a.sun.z * (p.z + a.R)) / r;
    return sun_visibility(a.A, a.T, a.tw, a.th, r, q, alt, mu_s) * a.A.solar;
}

// Environment term for a local-frame direction: the environment-luminance cube texel
// (AtmosphereRenderer.glsl:253-273) at kModelMatrix3 * d, from the sky-view LUT directly.
__device__ inline F3 environment(const PtArgs& a, F3 dl) {
    const F3 d = a.model0 * dl.x + a.model1 * dl.y + a.model2 * dl.z;
    return env_luminance(a.A, a.sky, a.T, a.tw, a.th, a.sv_l, a.sv_t, a.sw, a.sh, d);
}

__device__ __forceinline__ float pt_phase(const PtArgs& a, float c) {  // GetPhase (:93-96)
    float hb = henyey_greenstein(c, a.gb), hf = henyey_greenstein(c, a.gf);
    return hb * (1.0f - a.ratio) + hf * a.ratio;
}

// CreateOrthonormalBasis (shaders/Base/Common.glsl:44-51, Duff et al. 2017)
__device__ __forceinline__ void onb(F3 N, F3& t0, F3& t1) {
    float s = N.z >= 0.0f ? 1.0f : -1.0f;
    float aa = -1.0f / (s + N.z);
    float b = N.x * N.y * aa;
    t0 = {1.0f + s * N.x * N.x * aa, s * b, -s * N.x};
    t1 = {b, s + N.y * N.y * aa, -N.y};
}

// GenerateHGSample (:98-117)
__device__ inline F3 hg_sample(const PtArgs& a, Rng& rng, F3 d, float& value_over_pdf) {
    if (a.importance_sampling) {
        float g = rng.next() < a.ratio ? a.gf : a.gb;
        // HenyeyGreensteinInvertcdf (VolumetricCloudCommon.glsl:65-71)
        float xi = rng.next();
        float t = (1.0f - g * g) / (1.0f - g + 2.0f * g * xi);
        float cos_theta = (0.5f / g) * (1.0f + g * g - t * t);
        float sin_theta = sqrtf(clamp01(1.0f - cos_theta * cos_theta));
        F3 t0, t1;
        onb(d, t0, t1);
        float sp, cp;
        __sincosf(2.0f * kPiF * rng.next(), &sp, &cp);
        value_over_pdf = 1.0f;
        return t0 * (sin_theta * sp) + t1 * (sin_theta * cp) + d * cos_theta;
    }
    float sp, cp;  // UniformSphereSample (:50-55)
    __sincosf(2.0f * kPiF * rng.next(), &sp, &cp);
    float cos_theta = 1.0f - 2.0f * rng.next();
    float sin_theta = sqrtf(clamp01(1.0f - cos_theta * cos_theta));
    F3 dir = {cp * sin_theta, sp * sin_theta, cos_theta};
    value_over_pdf = pt_phase(a, dot(d, dir)) * (4.0f * kPiF);
    return dir;
}

// GenerateLambertSample (:119-129)
__device__ inline F3 lambert_sample(Rng& rng, F3 N) {
    float sin_theta = sqrtf(rng.next());
    float cos_theta = sqrtf(clamp01(1.0f - sin_theta * sin_theta));
    F3 t0, t1;
    onb(N, t0, t1);
    float sp, cp;
    __sincosf(2.0f * kPiF * rng.next(), &sp, &cp);
    return t0 * (sin_theta * sp) + t1 * (sin_theta * cp) + N * cos_theta;
}

// The kernel's arguments, read again from the kernarg segment (PtArgs is the kernel's only
// argument, at offset 0).  The empty asm hides that the pointer is loop-invariant, so each use
// after it issues its own scalar load instead of the whole block being loaded once and held in
// (spilled) SGPRs across the tracking loops.
__device__ __forceinline__ const PtArgs& kernel_args() {
    auto p = (const __attribute__((address_space(4))) PtArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const PtArgs*)p;
}

// One path between bounces: position, direction, throughput, radiance so far, and whether and
// where (distance from the camera) it first scattered.
struct PathState {
    F3 o, d, thr, L;
    float scattered_t;
    int istep;
    bool has_scattered;
};

// Trace's set-up (:160-181): the camera ray along `view`, clipped to the cloud region.  False
// when the path makes no bounce (it misses the region, or the bounce limit is 0).
__device__ __forceinline__ bool begin_path(const PtArgs& a, F3 view, PathState& p) {
    p.L = f3(0.0f);
    p.thr = f3(1.0f);
    p.has_scattered = false;
    p.scattered_t = 0.0f;
    p.istep = 0;
    p.o = a.cam;
    p.d = view;
    float c0, c1;
    region(a, p.o, p.d, c0, c1);
    if (c0 >= c1) return false;
    p.o = p.o + p.d * c0;
    return a.max_bounces > 0;
}

// One iteration of Trace's bounce loop (:182-247).  False when the path ends in it (it leaves
// the region, escapes, or reaches the bounce limit or zero throughput).
template <int kMat, bool kBricks, bool kCount, bool kCells, bool kFlat>
__device__ __forceinline__ bool bounce(Rng& rng, PathState& p, PtCount& wc) {
    const PtArgs& a = kernel_args();
    float t0, t1;
    region(a, p.o, p.d, t0, t1);
    if (t0 >= t1) return false;
    float t;
    if (!delta_track<kMat, kBricks, kCount, kCells, kFlat>(a, rng, p.o, p.d, t0, t1, t, wc)) {
        if (a.env == SKYHIP_ENV_OFF || !p.has_scattered) return false;
        if (a.env == SKYHIP_ENV_CONST_ENVIRONMENT_MAP) {
            p.L += p.thr * environment(a, p.d);
            return false;
        }
        const F3 o = p.o, d = p.d;
        float q = o.x * o.x + o.y * o.y + o.z * (o.z + 2.0f * a.R) + (a.R - a.A.rb) * (a.R + a.A.rb);
        float r = sqrtf(fmaxf(q + a.A.rb * a.A.rb, 0.0f));
        float alt = q / (r + a.A.rb);
        float mu = (d.x * o.x + d.y * o.y + d.z * (o.z + a.R)) / r;
        if (!ray_hits_ground(a.A, r, alt, mu)) {
            p.L += p.thr * environment(a, d);
            return false;
        }
        p.o = o + d * distance_to_bottom(a.A, r, alt, mu);
        F3 N = F3{p.o.x, p.o.y, p.o.z + a.R};
        N = N * (1.0f / length3(N));
        if (kCount) wc.vertices++;
        F3 bsdf = a.A.albedo * ((1.0f / kPiF) * dot(N, a.sun));
        p.L += p.thr * (ratio_track<kMat, kBricks, kCount, kCells, kFlat>(a, rng, p.o, a.sun, wc) * sun_illuminance(a, p.o) *
                        bsdf);
        if (a.env == SKYHIP_ENV_GROUND_SINGLE_BOUNCE) return false;
        p.d = lambert_sample(rng, N);
        p.thr *= a.A.albedo;
    } else {
        if (!p.has_scattered) p.scattered_t = length3(p.o - a.cam);  // the segment start, as :232-233
        p.has_scattered = true;
        p.o = p.o + p.d * t;
        if (kCount) wc.vertices++;
        // A real collision scatters with probability albedo (the reference's is 1): the sun
        // estimate at this vertex and the continuation both carry it.
        p.thr = p.thr * a.albedo;
        float ph = pt_phase(a, dot(p.d, a.sun));
        p.L += p.thr * (ratio_track<kMat, kBricks, kCount, kCells, kFlat>(a, rng, p.o, a.sun, wc) * ph * sun_illuminance(a, p.o));
        float vop;
        p.d = hg_sample(a, rng, p.d, vop);
        p.thr = p.thr * vop;
    }
    return ++p.istep < a.max_bounces && fmaxf(p.thr.x, fmaxf(p.thr.y, p.thr.z)) > 0.0f;
}

// Trace (:160-249)
template <int kMat, bool kBricks, bool kCount, bool kCells, bool kFlat>
__device__ F3 trace(Rng& rng, F3 view, bool& has_scattered, float& scattered_t, PtCount& wc) {
    PathState p;
    if (begin_path(kernel_args(), view, p))
        while (bounce<kMat, kBricks, kCount, kCells, kFlat>(rng, p, wc)) {
        }
    has_scattered = p.has_scattered;
    scattered_t = p.scattered_t;
    return p.L;
}

// A finished sample's aerial perspective (GetAerialPerspective, VolumetricCloudCommon.glsl:81-98,
// and the froxel visibility, :277) and its addition to the pixel's accumulator (:266-284).
// (u, v) is the pixel centre in [0, 1]^2 and view_z the z of its view direction.
__device__ __forceinline__ float4 add_sample(const PtArgs& a, float u, float v, float view_z, F3 L, bool hs, float st,
                                             float4 acc) {
    if (hs) {
        float r = a.cam.z + a.R, mu = view_z;
        float t = st;
        if (r > a.A.rt) {
            float disc = r * r * (mu * mu - 1.0f) + a.A.rt * a.A.rt;
            t = (mu < 0.0f && disc >= 0.0f) ? t - fmaxf(-r * mu - sqrtf(disc), 0.0f) : 0.0f;
        }
        float w = sqrtf(t / a.ap_max_distance);
        float fx = u * float(a.aw - 1), fy = v * float(a.ah - 1), fz = w * float(a.ad - 1);
        F3 apt = lut3d_trilinear(a.ap_t, a.aw, a.ah, a.ad, fx, fy, fz);
        F3 apl = lut3d_trilinear(a.ap_l, a.aw, a.ah, a.ad, fx, fy, fz);
        if (a.froxel) apl = apl * ray_scatter_visibility(a.froxel, a.fw, a.fh, a.fd, u, v, st, a.inv_froxel_max_distance);
        L = L * apt + apl;
    }
    acc.x += L.x;
    acc.y += L.y;
    acc.z += L.z;
    acc.w += hs ? 0.0f : 1.0f;
    return acc;
}

// Five waves per SIMD (96 VGPRs; 3 VGPRs of the path state spill to scratch): more waves hide
// more of the tracking's fetch latency.  Measured on the 1024^2 brick-majorant workload with the
// nested loops: 3 waves (no cap) 14.2 ms per frame, 4 waves 12.0, 5 waves 11.2, 6 waves 11.4, 7
// and 8 waves 12.0-12.2; with the flat loops and cell-corner lookups at 16^3: 4 waves 7.81-7.88,
// 5 waves 7.27-7.28 (24 VGPRs spilled), 6 waves 9.45; with the accumulator read after the path:
// 5 waves 7.26-7.27 (3 spilled), 6 waves 7.51-7.58 (21 spilled).  Every variant gives the same
// image.
template <int kMat, bool kBricks, bool kCount, bool kCells, bool kFlat>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void path_trace_kernel(PtArgs a) {
    const int lane = threadIdx.x;  // one wave64 per workgroup, an 8x8-pixel tile
    const int x = a.region[0] + blockIdx.x * 8 + (lane & 7);
    const int y = a.region[1] + blockIdx.y * 8 + (lane >> 3);
    PtCount wc;
    const bool active = x < a.region[2] && y < a.region[3];
    if (active) {
        F3 view;
        {
            const float u = (float(x) + 0.5f) / float(a.width), v = (float(y) + 0.5f) / float(a.height);
            float frag_dist;
            view = unproject(a.inv_mvp, u * 2.0f - 1.0f, v * 2.0f - 1.0f, 1.0f, a.cam, &frag_dist);
        }
        Rng rng;
        rng.pcg = a.prng;
        const uint32_t seed_xy = rng.hash(rng.hash(uint32_t(x)) + uint32_t(y));
        for (int f = 0; f < a.frames; ++f) {
            rng.s = rng.hash(seed_xy + a.first_frame + uint32_t(f));  // :260
            bool hs;
            float st;
            F3 L = trace<kMat, kBricks, kCount, kCells, kFlat>(rng, view, hs, st, wc);
            // The accumulator and the pixel centre are read again after the path (one read and
            // one write per frame, the same additions in frame order), so that they are not
            // held through the tracking loops.
            const PtArgs& ak = kernel_args();
            const float u = (float(x) + 0.5f) / float(ak.width), v = (float(y) + 0.5f) / float(ak.height);
            const long long pix = (long long)y * ak.width + x;
            ak.accum[pix] = add_sample(ak, u, v, view.z, L, hs, st, ak.accum[pix]);
        }
        if (a.mask) a.mask[(long long)y * a.width + x] = 1;
    }
    if (kCount) {
        unsigned long long v[5] = {active ? (unsigned long long)a.frames : 0ull, wc.lookups, wc.segments,
                                   wc.vertices, wc.shadow_rays};
        for (int i = 0; i < 5; ++i) {
            unsigned long long s = v[i];
            for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
            if (lane == 0) atomicAdd(&a.counters[i], s);
        }
    }
}

// The per-frame form of path_trace_kernel, persistent, with lanes refilled at every bounce.  The
// frame's pixels form 8x8 tiles; a wave takes tiles from a global counter and hands their pixels
// to its lanes.  After each bounce the lanes whose path has ended are found with a ballot, ranked
// with a population count of the lower lanes, and given the next pixels of the wave's tile (a new
// tile when it runs out), so a lane starts its next path while the rest of the wave is still
// tracking, where path_trace_kernel's lanes step through the bounces together and a wave lasts as
// long as its longest path.  Every path is the same computation with the same random sequence,
// and each pixel gets one path per launch, so the accumulator receives the same additions.  The
// loop ends when the counter has passed the last tile and no lane has a path.
template <int kMat, bool kBricks, bool kCells, bool kFlat>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void path_trace_refill_kernel(PtArgs) {
    const PtArgs& a0 = kernel_args();
    const int lane = threadIdx.x & 63;
    const int rx = a0.region[0], ry = a0.region[1], rx1 = a0.region[2], ry1 = a0.region[3];
    const int tiles_x = (rx1 - rx + 7) >> 3;
    const int n_tiles = tiles_x * ((ry1 - ry + 7) >> 3);
    Rng rng;
    rng.pcg = a0.prng;
    PathState p;
    PtCount wc;
    int px = 0, py = 0;
    float view_z = 0.0f;
    bool alive = false, exhausted = false;
    int next = 0, end = 0;  // the wave's unassigned pixels: items [next, end) of the frame (64 per tile)
    for (;;) {
        while (!exhausted) {  // wave-uniform: next, end and exhausted come from ballots and lane 0
            const unsigned long long idle = __ballot(!alive);
            if (idle == 0ull) break;
            if (next >= end) {
                int tile = 0;
                if (lane == 0) tile = int(atomicAdd(a0.queue, 1u));
                tile = __shfl(tile, 0);
                if (tile >= n_tiles) {
                    exhausted = true;
                    break;
                }
                next = tile * 64;
                end = next + 64;
            }
            const int avail = end - next;
            const int rank = __popcll(idle & ((1ull << lane) - 1ull));
            if (!alive && rank < avail) {
                const int item = next + rank, tile = item >> 6, j = item & 63;
                const int x = rx + (tile % tiles_x) * 8 + (j & 7), y = ry + (tile / tiles_x) * 8 + (j >> 3);
                if (x < rx1 && y < ry1) {
                    const PtArgs& a = kernel_args();
                    const float u = (float(x) + 0.5f) / float(a.width), v = (float(y) + 0.5f) / float(a.height);
                    float frag_dist;
                    const F3 view = unproject(a.inv_mvp, u * 2.0f - 1.0f, v * 2.0f - 1.0f, 1.0f, a.cam, &frag_dist);
                    rng.s = rng.hash(rng.hash(rng.hash(uint32_t(x)) + uint32_t(y)) + a.first_frame);  // :260
                    px = x;
                    py = y;
                    view_z = view.z;
                    alive = begin_path(a, view, p);
                    if (!alive) {
                        const long long pix = (long long)y * a.width + x;
                        a.accum[pix] = add_sample(a, u, v, view_z, p.L, p.has_scattered, p.scattered_t, a.accum[pix]);
                        if (a.mask) a.mask[pix] = 1;
                    }
                }
            }
            next += min(__popcll(idle), avail);
        }
        if (__ballot(alive) == 0ull) break;  // only once the tiles are exhausted
        if (alive) {
            alive = bounce<kMat, kBricks, false, kCells, kFlat>(rng, p, wc);
            if (!alive) {
                const PtArgs& a = kernel_args();
                const float u = (float(px) + 0.5f) / float(a.width), v = (float(py) + 0.5f) / float(a.height);
                const long long pix = (long long)py * a.width + px;
                a.accum[pix] = add_sample(a, u, v, view_z, p.L, p.has_scattered, p.scattered_t, a.accum[pix]);
                if (a.mask) a.mask[pix] = 1;
            }
        }
    }
}

__global__ void queue_reset_kernel(unsigned* q) {
    if (threadIdx.x == 0) q[0] = 0u;
}

// Display pass (:288-296).
__global__ __launch_bounds__(256) void path_trace_display_kernel(const float4* __restrict__ accum,
                                                                 const uint8_t* __restrict__ mask, uint32_t frame,
                                                                 uint16_t* __restrict__ display, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 s = accum[i];
    bool rendered = mask ? mask[i] != 0 : true;
    float inv = 1.0f / float(rendered ? frame : frame - 1u);
    float4 avg = make_float4(s.x * inv, s.y * inv, s.z * inv, s.w * inv);
    uint16_t* px = display + 4 * (long long)i;
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = half_bits_to_float(px[k]);
    float out[3] = {c[0] * avg.w + avg.x, c[1] * avg.w + avg.y, c[2] * avg.w + avg.z};
    for (int k = 0; k < 3; ++k) px[k] = __builtin_bit_cast(uint16_t, __float2half_rn(out[k]));
}

// Brick majorants of an FP16 grid: max over the B^3 brick plus a one-texel apron.
__global__ __launch_bounds__(256) void majorant_kernel(const uint16_t* __restrict__ g, int nx, int ny, int nz,
                                                       int bx, int by, int bz, int B, float* __restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= bx * by * bz) return;
    int x = i % bx, y = (i / bx) % by, z = i / (bx * by);
    int x0 = max(B * x - 1, 0), x1 = min(B * x + B, nx - 1);
    int y0 = max(B * y - 1, 0), y1 = min(B * y + B, ny - 1);
    int z0 = max(B * z - 1, 0), z1 = min(B * z + B, nz - 1);
    float m = 0.0f;
    for (int zz = z0; zz <= z1; ++zz)
        for (int yy = y0; yy <= y1; ++yy) {
            const uint16_t* row = g + ((long long)zz * ny + yy) * nx;
            for (int xx = x0; xx <= x1; ++xx) m = fmaxf(m, half_bits_to_float(row[xx]));
        }
    out[i] = m;
}

// The cell-corner copy: cell (i, j, k) from (-1, -1, -1), x fastest, gets the eight texels
// (i + a, j + b, k + c) in the order a + 2b + 4c, 0 outside the grid.
__global__ __launch_bounds__(256) void cells_kernel(const uint16_t* __restrict__ g, int nx, int ny, int nz,
                                                    uint4* __restrict__ out, unsigned n) {
    const unsigned c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const int cx = int(c % unsigned(nx + 1)) - 1, cy = int((c / unsigned(nx + 1)) % unsigned(ny + 1)) - 1,
              cz = int(c / (unsigned(nx + 1) * unsigned(ny + 1))) - 1;
    uint32_t h[8];
    for (int k = 0; k < 8; ++k) {
        const int x = cx + (k & 1), y = cy + ((k >> 1) & 1), z = cz + (k >> 2);
        const bool in = x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
        h[k] = in ? g[((long long)z * ny + y) * nx + x] : 0u;
    }
    out[c] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
}

// The walks' per-brick inverse majorants: the same float operations the walk would do per brick
// crossing (mu = majorant * scale, then 1 / mu when mu > 0), done once per brick.
__global__ __launch_bounds__(256) void inv_majorant_kernel(const float* __restrict__ majorants, float scale,
                                                           float* __restrict__ out, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float mu = majorants[i] * scale;
    out[i] = mu > 0.0f ? 1.0f / mu : 0.0f;
}

// Workgroups of path_trace_refill_kernel per CU: 5 of 4 waves, its occupancy at the 5-wave cap.
constexpr int kRefillBlocksPerCu = 5;
// Brick-majorant frames with the cell-corner copy run on one of two forms, chosen by the brick
// edge (same samples either way): the refill kernel for bricks of 4 texels, the per-pixel kernel
// above.  Measured on the 1024^2, 512^3-grid workload (ms per frame, one box): with one wave per
// 8x8-pixel workgroup the per-pixel kernel with flat tracking loops takes 6.65-6.66 at 16^3 and
// 7.74-7.77 at 8^3, against 8.38-8.48 for the refill kernel with nested loops at 8^3.  With
// 16x16-pixel workgroups the two forms were level at 8^3 (per-pixel flat 8.49-8.50, refill
// nested 8.43-8.48; refill flat 9.61-9.76, and 8.18-8.24 at 16^3): small bricks make crossings
// outnumber lookups (240 M against 82 M per frame at 8^3), and the flat loop carries the crossing
// body in every iteration.
constexpr int kRefillMaxBrick = 4;

template <bool kBricks, bool kCount>
static void launch_path_trace(hipStream_t st, const PtArgs& a, int cu_count) {
    // One wave64 per workgroup (8x8 pixels): a finished wave's slot takes the next tile by itself,
    // where a 256-lane workgroup's four waves are placed on one CU together.  Measured on the
    // 1024^2 workload: 6.71-6.73 against 7.26-7.30 ms per frame (16x16-pixel workgroups).
    dim3 grid(div_up(a.region[2] - a.region[0], 8), div_up(a.region[3] - a.region[1], 8));
    switch (a.mat.kind) {
        case SKYHIP_MATERIAL_DEFAULT0:
            hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_DEFAULT0, false, kCount, false, false>), grid, dim3(64), 0, st,
                               a);
            break;
        case SKYHIP_MATERIAL_DEFAULT1:
            hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_DEFAULT1, false, kCount, false, false>), grid, dim3(64), 0, st,
                               a);
            break;
        case SKYHIP_MATERIAL_DEFAULT1_WORLEY:
            hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_DEFAULT1_WORLEY, false, kCount, false, false>), grid, dim3(64), 0,
                               st, a);
            break;
        case SKYHIP_MATERIAL_MINIMAL:
            hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_MINIMAL, false, kCount, false, false>), grid, dim3(64), 0, st,
                               a);
            break;
        case SKYHIP_MATERIAL_VOXEL:
            hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_VOXEL, false, kCount, false, false>), grid, dim3(64), 0, st, a);
            break;
        default:
            if constexpr (kBricks && !kCount) {
                if (a.cells && a.brick <= kRefillMaxBrick) {
                    // One persistent launch per frame, its tile counter cleared by a kernel before it.
                    for (int f = 0; f < a.frames; ++f) {
                        PtArgs af = a;
                        af.first_frame = a.first_frame + uint32_t(f);
                        af.frames = 1;
                        hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, st, a.queue);
                        hipLaunchKernelGGL((path_trace_refill_kernel<SKYHIP_MATERIAL_VOXEL_F16, true, true, false>),
                                           dim3(cu_count * kRefillBlocksPerCu), dim3(256), 0, st, af);
                    }
                    break;
                }
            }
            if (kBricks && a.cells)
                hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_VOXEL_F16, kBricks, kCount, kBricks, kBricks>),
                                   grid, dim3(64), 0, st, a);
            else
                hipLaunchKernelGGL((path_trace_kernel<SKYHIP_MATERIAL_VOXEL_F16, kBricks, kCount, false, kBricks>),
                                   grid, dim3(64), 0, st, a);
            break;
    }
}

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_voxel_grid_brick_size(void) { return kDefaultBrick; }

static bool valid_brick(int b) { return b >= 4 && b <= 64 && (b & (b - 1)) == 0; }

extern "C" int skyhip_voxel_grid_majorants_sized(SkyhipContext* ctx, const uint16_t* grid, int32_t dim_x,
                                                 int32_t dim_y, int32_t dim_z, int32_t brick_size,
                                                 float* out_majorants) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!grid || !out_majorants) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dim_x < 2 || dim_y < 2 || dim_z < 2 || dim_x > 4096 || dim_y > 4096 || dim_z > 4096)
        return fail(SKYHIP_ERR_INVALID_SIZE, "grid dims outside [2, 4096]");
    if (!valid_brick(brick_size)) return fail(SKYHIP_ERR_INVALID_SIZE, "brick size must be a power of two in [4, 64]");
    const int B = brick_size;
    int bx = div_up(dim_x, B), by = div_up(dim_y, B), bz = div_up(dim_z, B);
    hipLaunchKernelGGL(majorant_kernel, dim3(div_up(bx * by * bz, 256)), dim3(256), 0, ctx->stream, grid, dim_x,
                       dim_y, dim_z, bx, by, bz, B, out_majorants);
    SKYHIP_LAUNCH_CHECK("majorant_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_voxel_grid_majorants(SkyhipContext* ctx, const uint16_t* grid, int32_t dim_x, int32_t dim_y,
                                           int32_t dim_z, float* out_majorants) {
    return skyhip_voxel_grid_majorants_sized(ctx, grid, dim_x, dim_y, dim_z, kDefaultBrick, out_majorants);
}

extern "C" int skyhip_voxel_grid_cells(SkyhipContext* ctx, const uint16_t* grid, int32_t dim_x, int32_t dim_y,
                                       int32_t dim_z, uint16_t* out_cells) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!grid || !out_cells) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dim_x < 2 || dim_y < 2 || dim_z < 2 || dim_x > 4096 || dim_y > 4096 || dim_z > 4096)
        return fail(SKYHIP_ERR_INVALID_SIZE, "grid dims outside [2, 4096]");
    if ((reinterpret_cast<uintptr_t>(out_cells) & 15) != 0)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "out_cells must be 16-byte aligned");
    const long long n = (long long)(dim_x + 1) * (dim_y + 1) * (dim_z + 1);
    if (n > (1ll << 31)) return fail(SKYHIP_ERR_INVALID_SIZE, "more than 2^31 cells");
    hipLaunchKernelGGL(cells_kernel, dim3(div_up(n, 256)), dim3(256), 0, ctx->stream, grid, dim_x, dim_y, dim_z,
                       reinterpret_cast<uint4*>(out_cells), (unsigned)n);
    SKYHIP_LAUNCH_CHECK("cells_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_path_trace(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                       const SkyhipAtmosphereRenderBufferData* render,
                                       const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                                       const SkyhipPathTracingParameters* params, const SkyhipPathTracingInputs* in,
                                       const int32_t region[4], uint32_t first_frame, int32_t frames,
                                       float* accumulation_rgba, uint8_t* rendered_mask, int32_t width,
                                       int32_t height, unsigned long long* work_counters) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!atmosphere || !render || !common || !params || !in || !region || !accumulation_rgba)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    if (!finite_block(reinterpret_cast<const float*>(common->uInvMVP), 16) || !finite_block(common->uCameraPos, 3) ||
        !finite_block(common->uSunDirection, 3) || !std::isfinite(common->uEarthRadius) ||
        !std::isfinite(common->uBottomAltitude) || !std::isfinite(common->uTopAltitude))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud common block has a non-finite field");
    if (!(common->uAerialPerspectiveLutMaxDistance > 0.0f))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "aerial-perspective max distance must be > 0");
    if (!finite_block(reinterpret_cast<const float*>(params), sizeof(*params) / 4))
        return fail(SKYHIP_ERR_NON_FINITE, "path-tracing parameters have a non-finite field");
    if (width < 1 || height < 1 || width > 16384 || height > 16384)
        return fail(SKYHIP_ERR_INVALID_SIZE, "image size outside [1, 16384]");
    if (region[0] < 0 || region[1] < 0 || region[2] > width || region[3] > height || region[0] > region[2] ||
        region[1] > region[3])
        return fail(SKYHIP_ERR_INVALID_SIZE, "render region outside the image");
    if (frames < 0 || frames > 65536) return fail(SKYHIP_ERR_INVALID_SIZE, "frames outside [0, 65536]");
    if (params->max_bounces < 0 || params->max_bounces > 4096)
        return fail(SKYHIP_ERR_INVALID_SIZE, "max_bounces outside [0, 4096]");
    if (params->prng != SKYHIP_PRNG_WANG_HASH && params->prng != SKYHIP_PRNG_PCG_HASH)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "unknown prng %d", params->prng);
    if (params->environment_lighting < SKYHIP_ENV_OFF || params->environment_lighting > SKYHIP_ENV_GROUND_MULTI_BOUNCE)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "unknown environment lighting %d", params->environment_lighting);
    if (params->importance_sampling &&
        (params->forward_phase_g == 0.0f || params->back_phase_g == 0.0f || fabsf(params->forward_phase_g) >= 1.0f ||
         fabsf(params->back_phase_g) >= 1.0f))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "HG importance sampling needs 0 < |g| < 1");
    if (!(params->region_box_half_width > 0.0f) || !(common->uTopAltitude > common->uBottomAltitude))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "empty cloud region");
    SKYHIP_CHECK(validate_lut(in->transmittance, in->transmittance_width, in->transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(in->sky_view_luminance, in->sky_view_width, in->sky_view_height, "sky-view luminance"));
    SKYHIP_CHECK(validate_lut(in->sky_view_transmittance, in->sky_view_width, in->sky_view_height,
                              "sky-view transmittance"));
    SKYHIP_CHECK(validate_lut(in->ap_luminance, in->ap_width, in->ap_height, "aerial-perspective luminance"));
    SKYHIP_CHECK(validate_lut(in->ap_transmittance, in->ap_width, in->ap_height, "aerial-perspective transmittance"));
    if (in->ap_depth < 2 || in->ap_depth > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "aerial-perspective depth");

    PtArgs a;
    std::memset(&a, 0, sizeof(a));
    SKYHIP_CHECK(build_material(material, a.mat));
    const bool bricks = in->brick_majorants != nullptr;
    if (bricks && a.mat.kind != SKYHIP_MATERIAL_VOXEL_F16)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "brick majorants need a VOXEL_F16 material");
    if (!bricks && !(params->sigma_t_max >= 0.0f)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "sigma_t_max < 0");
    if (!(params->albedo >= 0.0f)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "albedo < 0");

    a.A = make_atm(*atmosphere);
    a.cam = f3(common->uCameraPos);
    a.sun = f3(common->uSunDirection);
    a.R = common->uEarthRadius;
    a.hb = common->uBottomAltitude;
    a.ht = common->uTopAltitude;
    a.inv_thickness = 1.0f / (common->uTopAltitude - common->uBottomAltitude);
    a.half_width = params->region_box_half_width;
    std::memcpy(a.inv_mvp, common->uInvMVP, sizeof(a.inv_mvp));
    a.ap_max_distance = common->uAerialPerspectiveLutMaxDistance;
    a.max_bounces = params->max_bounces;
    a.importance_sampling = params->importance_sampling;
    a.prng = params->prng;
    a.env = params->environment_lighting;
    a.gf = params->forward_phase_g;
    a.gb = params->back_phase_g;
    a.ratio = params->forward_scattering_ratio;
    a.sigma_t_max = params->sigma_t_max;
    a.inv_sigma_t_max = params->sigma_t_max > 0.0f ? 1.0f / params->sigma_t_max : 0.0f;
    a.albedo = params->albedo;
    a.model0 = {params->model3[0], params->model3[1], params->model3[2]};
    a.model1 = {params->model3[3], params->model3[4], params->model3[5]};
    a.model2 = {params->model3[6], params->model3[7], params->model3[8]};
    a.sky = make_env_frame(*render, *atmosphere);
    a.T = reinterpret_cast<const float4*>(in->transmittance);
    a.tw = in->transmittance_width;
    a.th = in->transmittance_height;
    a.sv_l = reinterpret_cast<const float4*>(in->sky_view_luminance);
    a.sv_t = reinterpret_cast<const float4*>(in->sky_view_transmittance);
    a.sw = in->sky_view_width;
    a.sh = in->sky_view_height;
    a.ap_l = reinterpret_cast<const float4*>(in->ap_luminance);
    a.ap_t = reinterpret_cast<const float4*>(in->ap_transmittance);
    a.aw = in->ap_width;
    a.ah = in->ap_height;
    a.ad = in->ap_depth;
    a.froxel = ctx->froxel;
    a.fw = ctx->froxel_w, a.fh = ctx->froxel_h, a.fd = ctx->froxel_d;
    a.inv_froxel_max_distance = common->uInvShadowFroxelMaxDistance;
    if (a.froxel && !(a.inv_froxel_max_distance > 0.0f && std::isfinite(a.inv_froxel_max_distance)))
        return fail(SKYHIP_ERR_NON_FINITE, "a shadow froxel is bound but uInvShadowFroxelMaxDistance is not > 0");
    if (bricks) {
        const SkyhipCloudVoxelBufferData& vb = material->voxel;
        a.majorants = in->brick_majorants;
        a.brick = in->brick_size == 0 ? kDefaultBrick : in->brick_size;
        if (!valid_brick(a.brick)) return fail(SKYHIP_ERR_INVALID_SIZE, "brick size must be 0 or a power of two in [4, 64]");
        a.inv_brick = 1.0f / float(a.brick);
        for (int k = 0; k < 3; ++k) {
            a.nb[k] = div_up(material->voxel_grid_dim[k], a.brick);
            a.grid_n[k] = float(material->voxel_grid_dim[k]);
        }
        // X = (pos * freq + bias) * N ; Z = (z - hb) / (ht - hb) * N (BrickWalk subtracts hb)
        a.tex_scale[0] = vb.uSampleFrequency[0] * a.grid_n[0];
        a.tex_bias[0] = vb.uSampleBias[0] * a.grid_n[0];
        a.tex_scale[1] = vb.uSampleFrequency[1] * a.grid_n[1];
        a.tex_bias[1] = vb.uSampleBias[1] * a.grid_n[1];
        a.tex_scale[2] = a.inv_thickness * a.grid_n[2];
        a.tex_bias[2] = 0.0f;
        a.majorant_scale = vb.uDensity;
        a.cells = reinterpret_cast<const uint4*>(in->cell_corners);
        if (a.cells && (reinterpret_cast<uintptr_t>(a.cells) & 15) != 0)
            return fail(SKYHIP_ERR_INVALID_ARGUMENT, "cell_corners must be 16-byte aligned");
        if (a.cells) {  // a copy has at most 2^31 cells (skyhip_voxel_grid_cells); the lookups index in 32 bits
            const long long cells = (material->voxel_grid_dim[0] + 1LL) * (material->voxel_grid_dim[1] + 1LL) *
                                    (material->voxel_grid_dim[2] + 1LL);
            if (cells > (1LL << 31))
                return fail(SKYHIP_ERR_INVALID_SIZE, "cell_corners: a grid of more than 2^31 cells has no copy");
        }
        if (!(vb.uDensity >= 0.0f)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "voxel density < 0");
    }
    for (int k = 0; k < 4; ++k) a.region[k] = region[k];
    a.first_frame = first_frame;
    a.frames = frames;
    a.width = width;
    a.height = height;
    a.accum = reinterpret_cast<float4*>(accumulation_rgba);
    a.mask = rendered_mask;
    a.counters = work_counters;
    if (region[2] == region[0] || region[3] == region[1] || frames == 0) return SKYHIP_OK;
    if (!ctx->pt_queue) SKYHIP_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->pt_queue), 256));
    a.queue = ctx->pt_queue;
    if (bricks) {
        const size_t n = size_t(a.nb[0]) * a.nb[1] * a.nb[2];
        if (ctx->pt_inv_majorants_count < n) {
            if (ctx->pt_inv_majorants) {
                SKYHIP_HIP(hipStreamSynchronize(ctx->stream));
                SKYHIP_HIP(hipFree(ctx->pt_inv_majorants));
                ctx->pt_inv_majorants = nullptr;
                ctx->pt_inv_majorants_count = 0;
            }
            SKYHIP_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->pt_inv_majorants), n * sizeof(float)));
            ctx->pt_inv_majorants_count = n;
        }
        hipLaunchKernelGGL(inv_majorant_kernel, dim3(div_up((long long)n, 256)), dim3(256), 0, ctx->stream,
                           a.majorants, a.majorant_scale, ctx->pt_inv_majorants, int(n));
        SKYHIP_LAUNCH_CHECK("inv_majorant_kernel");
        a.inv_majorants = ctx->pt_inv_majorants;
    }
    const int cus = std::max(ctx->cu_count, 1);
    if (bricks) {
        if (work_counters)
            launch_path_trace<true, true>(ctx->stream, a, cus);
        else
            launch_path_trace<true, false>(ctx->stream, a, cus);
    } else {
        if (work_counters)
            launch_path_trace<false, true>(ctx->stream, a, cus);
        else
            launch_path_trace<false, false>(ctx->stream, a, cus);
    }
    SKYHIP_LAUNCH_CHECK("path_trace_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_path_trace_display(SkyhipContext* ctx, const float* accumulation_rgba,
                                               const uint8_t* rendered_mask, uint32_t frame_id,
                                               void* display_rgba16f, int32_t width, int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!accumulation_rgba || !display_rgba16f) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (width < 1 || height < 1 || width > 16384 || height > 16384)
        return fail(SKYHIP_ERR_INVALID_SIZE, "image size outside [1, 16384]");
    if (frame_id < 1) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "frame_id starts at 1");
    int n = width * height;
    hipLaunchKernelGGL(path_trace_display_kernel, dim3(div_up(n, 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const float4*>(accumulation_rgba), rendered_mask, frame_id,
                       reinterpret_cast<uint16_t*>(display_rgba16f), n);
    SKYHIP_LAUNCH_CHECK("path_trace_display_kernel");
    return SKYHIP_OK;
}
