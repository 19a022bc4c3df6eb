// Device-side atmosphere model shared by the LUT kernels, the batched LUT
// builder, the surrogate's label generator and the cloud passes.
//
// The arithmetic restates shaders/SkyRendering/Atmosphere.glsl and
// AtmosphereRenderer.glsl in forms that avoid FP32 cancellation at planetary
// radii, with results equal to the GLSL expressions in exact arithmetic:
//   r_d^2 - R_b^2   = d (d + 2 r mu) + (r - R_b)(r + R_b)        (altitude without r_d - R_b)
//   altitude_d      = (r_d^2 - R_b^2) / (r_d + R_b)
//   rho_d           = sqrt(r_d^2 - R_b^2)                          (GetTransmittanceTextureUvFromRMu)
//   R_t^2 - r_d^2   = H^2 - rho_d^2                                (DistanceToTopAtmosphereBoundary)
//   mu_s at d       = (r mu_s0 + d nu) / r_d                      (dot(sun, normalize(p_d - c)))
//   (L - L T_i)/s   = L (1 - T_i)/s, 1 - T_i = -expm1(-s dx)      (Atmosphere.glsl:288,290)
//   texel coord     = x (N - 1)                                    (u N - 0.5 with u = 0.5/N + x(1 - 1/N))
// Texture fetches are software bilinear with exact weights (GL LINEAR,
// CLAMP_TO_EDGE), not the hardware sampler, whose fixed-point weights would
// not meet the 1e-3 relative bar.
#pragma once

#include "common.h"

namespace skyhip {

constexpr float kPiF = 3.14159265358979323846f;

struct AtmDev {
    F3 solar;
    float sun_angular_radius;
    F3 rayleigh;
    float inv_hr;
    F3 mie_s;
    float inv_hm;
    F3 mie_ext;  // mie scattering + mie absorption
    float ozone_center;
    F3 ozone;
    float inv_ozone_width;
    F3 albedo;
    float g;
    float ms_mask;
    float rb, rt;
    float thickness;  // rt - rb
    float H2;         // rt^2 - rb^2
    float H;
    float inv_H;
    float inv_thickness;
    float t_steps, ms_steps;
};

__host__ __device__ inline AtmDev make_atm(const SkyhipAtmosphereBufferData& b) {
    AtmDev a;
    a.solar = f3(b.solar_illuminance);
    a.sun_angular_radius = b.sun_angular_radius;
    a.rayleigh = f3(b.rayleigh_scattering);
    a.inv_hr = b.inv_rayleigh_exponential_distribution;
    a.mie_s = f3(b.mie_scattering);
    a.inv_hm = b.inv_mie_exponential_distribution;
    a.mie_ext = f3(b.mie_scattering) + f3(b.mie_absorption);
    a.ozone_center = b.ozone_center_altitude;
    a.ozone = f3(b.ozone_absorption);
    a.inv_ozone_width = b.inv_ozone_width;
    a.albedo = f3(b.ground_albedo);
    a.g = b.mie_phase_g;
    a.ms_mask = b.multiscattering_mask;
    a.rb = b.bottom_radius;
    a.rt = b.top_radius;
    a.thickness = b.top_radius - b.bottom_radius;
    a.H2 = a.thickness * (b.top_radius + b.bottom_radius);
    a.H = sqrtf(a.H2);
    a.inv_H = 1.0f / a.H;
    a.inv_thickness = 1.0f / a.thickness;
    a.t_steps = b.transmittance_steps;
    a.ms_steps = b.multiscattering_steps;
    return a;
}

// Division and square root through the hardware's reciprocal and square-root instructions
// (about 1 ulp) for the integrators' per-step arithmetic; the correctly rounded forms expand
// to about ten instructions each, and a step has nine divisions and three square roots.
__device__ __forceinline__ float fdiv_fast(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

// 1 - exp(-x) for x >= 0 given e = exp(-x): below x = 1/8 the series x - x^2/2 + ... + x^5/120
// (truncation < 1e-8 relative), which keeps the relative precision that 1 - e loses to
// cancellation for small x; above, 1 - e (cancellation costs at most 1/(1 - e^-1/8) ~ 8.5x
// of e's rounding).
__device__ __forceinline__ float one_minus_exp(float x, float e) {
    const float series = x * (1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f - x * (1.0f / 120.0f)))));
    return x < 0.125f ? series : 1.0f - e;
}

// Atmosphere.glsl:119-132 / :156-159 split into densities so the exps are shared.
struct Densities {
    float rayleigh, mie, ozone;
};
__device__ inline Densities densities(const AtmDev& A, float altitude) {
    Densities d;
    d.rayleigh = clamp01(__expf(-altitude * A.inv_hr));
    d.mie = clamp01(__expf(-altitude * A.inv_hm));
    d.ozone = fmaxf(0.0f, 1.0f - fabsf(altitude - A.ozone_center) * A.inv_ozone_width);
    return d;
}
__device__ inline F3 extinction(const AtmDev& A, const Densities& d) {
    return A.rayleigh * d.rayleigh + A.mie_ext * d.mie + A.ozone * d.ozone;
}

// Atmosphere.glsl:138-154 (non-multiscattering variants).
__device__ inline float rayleigh_phase(float c) { return 3.0f / (16.0f * kPiF) * (1.0f + c * c); }
__device__ inline float mie_phase(float g, float c) {
    float k = 3.0f / (8.0f * kPiF) * (1.0f - g * g) / (2.0f + g * g);
    float b = 1.0f + g * g - 2.0f * g * c;
    return k * (1.0f + c * c) / (b * sqrtf(b));
}

// Bilinear CLAMP_TO_EDGE fetch of RGB from an RGBA32F LUT at texel-space
// coordinates (0 .. N-1 covers texel centres 0 .. N-1).  kCells: `lut` is the LUT's cell-corner
// copy (lut_cells_kernel), where texel (x, y) is followed by the rest of its bilinear footprint:
// 4 float4 per texel, (x, y), (x1, y), (x, y1), (x1, y1) with x1 = min(x + 1, w - 1) and y1 likewise,
// read from one address.  The same texels and the same arithmetic either way.
template <bool kCells = false>
__device__ inline F3 lut_bilinear(const float4* __restrict__ lut, int w, int h, float fx, float fy) {
    fx = fminf(fmaxf(fx, 0.0f), float(w - 1));
    fy = fminf(fmaxf(fy, 0.0f), float(h - 1));
    int x0 = int(fx), y0 = int(fy);
    float ax = fx - float(x0), ay = fy - float(y0);
    float4 a, b, c, d;
    if (kCells) {
        const float4* q = lut + 4 * (y0 * w + x0);
        a = q[0], b = q[1], c = q[2], d = q[3];
    } else {
        int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
        a = lut[y0 * w + x0], b = lut[y0 * w + x1], c = lut[y1 * w + x0], d = lut[y1 * w + x1];
    }
    float bx = 1.0f - ax, by = 1.0f - ay;
    return {by * (bx * a.x + ax * b.x) + ay * (bx * c.x + ax * d.x),
            by * (bx * a.y + ax * b.y) + ay * (bx * c.y + ax * d.y),
            by * (bx * a.z + ax * b.z) + ay * (bx * c.z + ax * d.z)};
}

// GetSunVisibility (Atmosphere.glsl:105-117) at a point given by its radius r,
// q = r^2 - R_b^2 (negative below the ground, where the reference's aerial-
// perspective froxels can reach), altitude and mu_s.
template <bool kCells = false>
__device__ inline F3 sun_visibility(const AtmDev& A, const float4* __restrict__ T, int tw, int th, float r,
                                    float q, float altitude, float mu_s) {
    // GetTransmittanceTextureUvFromRMu (:90-103): rho = SafeSqrt(r^2 - R_b^2),
    // discriminant r^2 (mu^2 - 1) + R_t^2 = r^2 mu^2 + (H^2 - q).
    float rho = fsqrt_fast(fmaxf(q, 0.0f));
    float disc = r * r * mu_s * mu_s + (A.H2 - q);
    float d = fmaxf(-r * mu_s + fsqrt_fast(fmaxf(disc, 0.0f)), 0.0f);
    float d_min = A.thickness - altitude;
    float d_max = rho + A.H;
    float x_mu = fdiv_fast(d - d_min, d_max - d_min);
    float x_r = rho * A.inv_H;
    F3 t = lut_bilinear<kCells>(T, tw, th, x_mu * float(tw - 1), x_r * float(th - 1));
    const float inv_r = __builtin_amdgcn_rcpf(r);
    float sin_theta_h = A.rb * inv_r;
    float cos_theta_h = -rho * inv_r;
    float e = sin_theta_h * A.sun_angular_radius;
    // smoothstep(-e, e, x)
    float s = clamp01(fdiv_fast(mu_s - cos_theta_h + e, 2.0f * e));
    return t * (s * s * (3.0f - 2.0f * s));
}

// GetMultiscatteringContribution (Atmosphere.glsl:169-178).
template <bool kCells = false>
__device__ inline F3 ms_contribution(const AtmDev& A, const float4* __restrict__ MS, int mw, int mh, float altitude,
                                     float mu_s) {
    float x_mu_s = mu_s * 0.5f + 0.5f;
    float x_r = altitude * A.inv_thickness;
    return lut_bilinear<kCells>(MS, mw, mh, x_mu_s * float(mw - 1), x_r * float(mh - 1));
}

// Geometry of a ray that starts at radius r (altitude alt0 = r - R_b) with
// zenith cosine mu, sun zenith cosine mu_s0 and sun-view cosine nu.
struct RayStart {
    float r, alt0, mu, mu_s0, nu;
};

// RayIntersectsGround / DistanceTo{Top,Bottom}AtmosphereBoundary (Atmosphere.glsl:57-69)
__device__ inline bool ray_hits_ground(const AtmDev& A, float r, float alt0, float mu) {
    return mu < 0.0f && r * r * mu * mu - alt0 * (r + A.rb) >= 0.0f;
}
// Roots of t^2 + 2 r mu t - c = 0 in the conjugate form that does not cancel:
// the far root -r mu + sqrt(disc) equals c / (r mu + sqrt(disc)) when mu > 0,
// and the near root -r mu - sqrt(disc) equals -c / (-r mu + sqrt(disc)).
__device__ inline float distance_to_top(const AtmDev& A, float r, float alt0, float mu) {
    float c = A.H2 - alt0 * (r + A.rb);  // R_t^2 - r^2
    float disc = r * r * mu * mu + c;
    float sq = sqrtf(fmaxf(disc, 0.0f));
    float d = mu > 0.0f ? c / (r * mu + sq) : -r * mu + sq;
    return fmaxf(d, 0.0f);
}
__device__ inline float distance_to_bottom(const AtmDev& A, float r, float alt0, float mu) {
    float c = alt0 * (r + A.rb);  // r^2 - R_b^2
    float disc = r * r * mu * mu - c;
    float sq = sqrtf(fmaxf(disc, 0.0f));
    float d = (mu < 0.0f && disc >= 0.0f) ? c / (-r * mu + sq) : -r * mu - sq;
    return fmaxf(d, 0.0f);
}

// Per-ray constants of the integrator below.
struct ScatterRay {
    float dx, rp, mp, two_r_mu, rho0_2, rb2;
};
template <bool kMS>
__device__ inline ScatterRay scatter_ray(const AtmDev& A, const RayStart& s, float marching_distance, float steps) {
    const float iso = 1.0f / (4.0f * kPiF);
    return {marching_distance / steps, kMS ? iso : rayleigh_phase(s.nu), kMS ? iso : mie_phase(A.g, s.nu),
            2.0f * s.r * s.mu, s.alt0 * (s.r + A.rb), A.rb * A.rb};
}

// One step of the integrator at float counter value i (kCells: T and MS are cell-corner copies).
template <bool kMS, bool kCells = false>
__device__ __forceinline__ void scatter_step(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                             const float4* __restrict__ MS, int mw, int mh, const RayStart& s,
                                             const ScatterRay& c, float i, F3& L, F3& Tr, F3& Lf) {
    const float dx = c.dx;
    float d = i * dx;
    float q = fmaf(d, d + c.two_r_mu, c.rho0_2);  // r_d^2 - R_b^2, negative below ground
    float r_d = fsqrt_fast(fmaxf(q + c.rb2, 0.0f));
    float alt = fdiv_fast(q, r_d + A.rb);
    float mu_s = fdiv_fast(s.r * s.mu_s0 + d * s.nu, r_d);
    Densities dn = densities(A, alt);
    F3 sr = A.rayleigh * dn.rayleigh;
    F3 sm = A.mie_s * dn.mie;
    F3 ext = sr + A.mie_ext * dn.mie + A.ozone * dn.ozone;
    const F3 tr = {__expf(-ext.x * dx), __expf(-ext.y * dx), __expf(-ext.z * dx)};
    F3 one_minus_t = {one_minus_exp(ext.x * dx, tr.x), one_minus_exp(ext.y * dx, tr.y),
                      one_minus_exp(ext.z * dx, tr.z)};
    F3 Li = (sr * c.rp + sm * c.mp) * sun_visibility<kCells>(A, T, tw, th, r_d, q, alt, mu_s);
    F3 inv_ext = {__builtin_amdgcn_rcpf(ext.x), __builtin_amdgcn_rcpf(ext.y), __builtin_amdgcn_rcpf(ext.z)};
    if (!kMS) {
        Li += A.ms_mask * ms_contribution<kCells>(A, MS, mw, mh, alt, mu_s) * (sr + sm);
        Li *= A.solar;
    }
    L += Tr * Li * one_minus_t * inv_ext;
    if (kMS) Lf += Tr * (sr + sm) * one_minus_t * inv_ext;
    Tr *= tr;
}

// ComputeScatteredLuminance (Atmosphere.glsl:220-295).  kMS selects the
// MULTISCATTERING_COMPUTE_PROGRAM variant.  The loop keeps the shader's float
// counter (i = start_i; i < steps; ++i), so fractional step counts iterate
// exactly as the GLSL does.
template <bool kMS, bool kCells = false>
__device__ inline void scattered_luminance(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                           const float4* __restrict__ MS, int mw, int mh, const RayStart& s,
                                           float marching_distance, float steps, float start_i, F3& L, F3& Tr,
                                           F3& Lf) {
    const ScatterRay c = scatter_ray<kMS>(A, s, marching_distance, steps);
    L = f3(0.0f);
    Tr = f3(1.0f);
    Lf = f3(0.0f);
    for (float i = kMS ? 0.5f : start_i; i < steps; i += 1.0f)
        scatter_step<kMS, kCells>(A, T, tw, th, MS, mw, mh, s, c, i, L, Tr, Lf);
}

// `count` consecutive steps of the same float counter, starting at counter value i0, with
// (L, Tr) starting at (0, 1): a contiguous part of scattered_luminance's sum.  Parts combine
// in order as L = L_a + Tr_a L_b, Tr = Tr_a Tr_b (exact in real arithmetic).
__device__ inline void scattered_luminance_part(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                                const float4* __restrict__ MS, int mw, int mh, const RayStart& s,
                                                const ScatterRay& c, float i0, int count, F3& L, F3& Tr) {
    L = f3(0.0f);
    Tr = f3(1.0f);
    F3 Lf = f3(0.0f);
    float i = i0;
    for (int k = 0; k < count; ++k, i += 1.0f) scatter_step<false>(A, T, tw, th, MS, mw, mh, s, c, i, L, Tr, Lf);
}

// The same for the multiple-scattering variant (isotropic phase, no MS term, and the
// multiple-scattering transfer Lf beside L): parts combine as L = L_a + Tr_a L_b,
// Lf = Lf_a + Tr_a Lf_b, Tr = Tr_a Tr_b.
__device__ inline void scattered_luminance_part_ms(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                                   const RayStart& s, const ScatterRay& c, float i0, int count, F3& L,
                                                   F3& Tr, F3& Lf) {
    L = f3(0.0f);
    Tr = f3(1.0f);
    Lf = f3(0.0f);
    float i = i0;
    for (int k = 0; k < count; ++k, i += 1.0f) scatter_step<true>(A, T, tw, th, nullptr, 0, 0, s, c, i, L, Tr, Lf);
}

// scattered_luminance<false> spread over kParts consecutive lanes of a wave (kParts a power of
// two dividing 64; part = lane index within the group).  Part p integrates steps
// [p c, (p + 1) c), c = ceil(n / kParts), of the shader's float counter (n its trip count,
// i0 the counter value the loop reaches at step p c), and part 0 combines the parts in order:
// L = L0 + Tr0 (L1 + Tr1 (...)), Tr = Tr0 Tr1 ...  All lanes of a group must call this
// together (it shuffles); `active` (marching_distance > 0) must be the same across a group.
// The result is valid on part 0.
template <int kParts>
__device__ inline void scattered_luminance_split(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                                 const float4* __restrict__ MS, int mw, int mh, const RayStart& s,
                                                 float marching_distance, float steps, float start_i, bool active,
                                                 int part, F3& L, F3& Tr) {
    L = f3(0.0f);
    Tr = f3(1.0f);
    if (active) {
        int n = 0;
        for (float i = start_i; i < steps; i += 1.0f) ++n;
        const int c = (n + kParts - 1) / kParts, first = part * c;
        float i0 = start_i;
        for (int k = 0; k < first && k < n; ++k) i0 += 1.0f;
        const ScatterRay sc = scatter_ray<false>(A, s, marching_distance, steps);
        scattered_luminance_part(A, T, tw, th, MS, mw, mh, s, sc, i0, max(0, min(c, n - first)), L, Tr);
    }
    F3 accL = f3(0.0f), accT = f3(1.0f);  // the parts after this lane's, combined
    for (int p = kParts - 1; p >= 1; --p) {
        const F3 Lp = {__shfl_down(L.x, p, 64), __shfl_down(L.y, p, 64), __shfl_down(L.z, p, 64)};
        const F3 Tp = {__shfl_down(Tr.x, p, 64), __shfl_down(Tr.y, p, 64), __shfl_down(Tr.z, p, 64)};
        if (p == kParts - 1) {
            accL = Lp;
            accT = Tp;
        } else {
            accL = Lp + Tp * accL;
            accT = Tp * accT;
        }
    }
    if (kParts > 1) {
        L = L + Tr * accL;
        Tr = Tr * accT;
    }
}

}  // namespace skyhip
