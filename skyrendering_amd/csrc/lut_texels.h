// Per-texel bodies of the four atmosphere LUT passes, shared by the single-set
// kernels (atmosphere.hip) and the batched builder (batch.hip), so a batch item
// runs exactly the single-set arithmetic.
//
//   transmittance_texel   Atmosphere.glsl:71-88, :311-337 (FP32, or FP64 near the horizon)
//   multiscattering_texel Atmosphere.glsl:344-438, one wave64 per texel
//   sky_view_texel        AtmosphereRenderer.glsl:56-186
//   aerial_ray            AtmosphereRenderer.glsl:191-243 (set-up; the integration is split over lanes)
#pragma once

#include "atmosphere_device.h"

namespace skyhip {

__device__ inline float sexp(float x) { return __expf(x); }
__device__ inline double sexp(double x) { return exp(x); }
__device__ inline float ssqrt(float x) { return sqrtf(x); }
__device__ inline double ssqrt(double x) { return sqrt(x); }

// The density integrals (Rayleigh, Mie, ozone) over steps [first, first + count) of one
// transmittance texel's march (the shader's float counter runs 0.5, 1.5, ...), evaluated
// entirely in scalar type S, and the step length dx.  Extinction is linear in the integrals,
// so parts of the march add.
template <class S>
__device__ inline void transmittance_od(const SkyhipAtmosphereBufferData& b, int ix, int iy, int w, int h, int first,
                                        int count, S od[3], S& dx_out) {
    const S rb = b.bottom_radius, rt = b.top_radius;
    const S thickness = rt - rb;
    const S H2 = thickness * (rt + rb);
    const S H = ssqrt(H2);
    const S x_mu = S(ix) / S(w - 1);
    const S x_r = S(iy) / S(h - 1);
    // GetRMuFromTransmittanceTextureIndex
    const S rho = H * x_r;
    const S rho2 = rho * rho;
    const S r = ssqrt(rho2 + rb * rb);
    const S alt0 = rho2 / (r + rb);  // r - R_b
    const S d_min = thickness - alt0;
    const S d_max = rho + H;
    const S d = d_min + x_mu * (d_max - d_min);
    S mu = d == S(0) ? S(1) : ((H - rho) * (H + rho) - d * d) / (S(2) * r * d);
    mu = mu < S(-1) ? S(-1) : (mu > S(1) ? S(1) : mu);
    // ComputeTransmittanceToTopAtmosphereBoundary
    const S disc = r * r * mu * mu + (H2 - rho2);
    S dist = -r * mu + ssqrt(disc > S(0) ? disc : S(0));
    dist = dist > S(0) ? dist : S(0);
    const float steps = b.transmittance_steps;
    const S dx = dist / S(steps);
    const S two_r_mu = S(2) * r * mu;
    const S rb2 = rb * rb;
    const S inv_hr = b.inv_rayleigh_exponential_distribution, inv_hm = b.inv_mie_exponential_distribution;
    const S oc = b.ozone_center_altitude, inv_ow = b.inv_ozone_width;
    S od_r = 0, od_m = 0, od_o = 0;
    float i = 0.5f + float(first);  // exact: the counter's values are k + 0.5 far below 2^23
    for (int k = 0; k < count; ++k, i += 1.0f) {
        S di = S(i) * dx;
        S q = di * (di + two_r_mu) + rho2;  // r_i^2 - R_b^2
        S rd = ssqrt(q + rb2);
        S alt = q / (rd + rb);
        S er = sexp(-alt * inv_hr), em = sexp(-alt * inv_hm);
        er = er > S(1) ? S(1) : er;
        em = em > S(1) ? S(1) : em;
        S dev = alt - oc;
        S eo = S(1) - (dev < S(0) ? -dev : dev) * inv_ow;
        eo = eo > S(0) ? eo : S(0);
        od_r += er;
        od_m += em;
        od_o += eo;
    }
    od[0] = od_r;
    od[1] = od_m;
    od[2] = od_o;
    dx_out = dx;
}

// Trip count of the transmittance march's float counter (0.5, 1.5, ... < steps).
__device__ inline int transmittance_step_count(const SkyhipAtmosphereBufferData& b) {
    int n = 0;
    for (float i = 0.5f; i < b.transmittance_steps; i += 1.0f) ++n;
    return n;
}

template <class S>
__device__ inline F3 transmittance_from_od(const SkyhipAtmosphereBufferData& b, const S od[3], S dx) {
    F3 out;
    for (int c = 0; c < 3; ++c) {
        S v = (S(b.rayleigh_scattering[c]) * od[0] + (S(b.mie_scattering[c]) + S(b.mie_absorption[c])) * od[1] +
               S(b.ozone_absorption[c]) * od[2]) * dx;
        (&out.x)[c] = float(sexp(-v));
    }
    return out;
}

// One transmittance texel, evaluated entirely in scalar type S.
template <class S>
__device__ inline F3 transmittance_texel(const SkyhipAtmosphereBufferData& b, int ix, int iy, int w, int h) {
    S od[3], dx;
    transmittance_od<S>(b, ix, iy, w, h, 0, transmittance_step_count(b), od, dx);
    return transmittance_from_od<S>(b, od, dx);
}

// Texels with x >= this column (the near-horizon end of the x_mu axis, where the
// optical depth is largest) integrate in FP64.
__host__ __device__ inline int transmittance_fp64_column(int w) { return (w * 3) / 4; }

__device__ inline float4 transmittance_texel_rgba(const SkyhipAtmosphereBufferData& b, int ix, int iy, int w,
                                                  int h) {
    F3 v = ix >= transmittance_fp64_column(w) ? transmittance_texel<double>(b, ix, iy, w, h)
                                              : transmittance_texel<float>(b, ix, iy, w, h);
    return make_float4(v.x, v.y, v.z, 1.0f);
}

// The ray of one multiple-scattering direction (lane = sphere direction index, 0..63).
struct MsRay {
    RayStart s;
    float dist;
    bool hit;
};
__device__ inline MsRay ms_ray(const AtmDev& A, int ix, int iy, int w, int h, int lane) {
    // GetAltitudeMuSFromMultiscatteringTextureIndex
    const float altitude = (float(iy) / float(h - 1)) * A.thickness;
    const float mu_s = (float(ix) / float(w - 1)) * 2.0f - 1.0f;
    const float sun_z = sqrtf(1.0f - mu_s * mu_s);
    // GetDirectionFromLocalIndex
    const float unit_theta = (0.5f + float(lane >> 3)) / 8.0f;
    const float unit_phi = (0.5f + float(lane & 7)) / 8.0f;
    const float cos_theta = 1.0f - 2.0f * unit_theta;
    const float sin_theta = sqrtf(clamp01(1.0f - cos_theta * cos_theta));
    float sin_phi, cos_phi;
    sincosf(2.0f * kPiF * unit_phi, &sin_phi, &cos_phi);
    const F3 view = {cos_phi * sin_theta, cos_theta, sin_phi * sin_theta};

    const float r = altitude + A.rb;
    const float mu = view.y;
    const float nu = view.y * mu_s + view.z * sun_z;
    const bool hit = ray_hits_ground(A, r, altitude, mu);
    const float dist = hit ? distance_to_bottom(A, r, altitude, mu) : distance_to_top(A, r, altitude, mu);
    return {RayStart{r, altitude, mu, mu_s, nu}, dist, hit};
}

// The ground term (ComputeGroundLuminance, :297-306, at |p - c| = R_b) and the reduction over the
// 64 directions of a wave, in the reference's fixed-order shared-memory tree (:392-427): lane i
// adds lane i + s for s = 32, 16, ..., 1, which is exactly this shuffle chain.  Valid in lane 0.
template <bool kCells = false>
__device__ inline float4 multiscattering_reduce(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                                const MsRay& ray, F3 L, F3 Tr, F3 Lf) {
    if (ray.hit) {
        float mu_g = (ray.s.r * ray.s.mu_s0 + ray.dist * ray.s.nu) / A.rb;
        F3 sv = sun_visibility<kCells>(A, T, tw, th, A.rb, 0.0f, 0.0f, mu_g);
        L += Tr * ((1.0f / kPiF) * clamp01(mu_g)) * A.albedo * sv;
    }
    for (int s = 32; s >= 1; s >>= 1) {
        L.x += __shfl_down(L.x, s, 64);
        L.y += __shfl_down(L.y, s, 64);
        L.z += __shfl_down(L.z, s, 64);
        Lf.x += __shfl_down(Lf.x, s, 64);
        Lf.y += __shfl_down(Lf.y, s, 64);
        Lf.z += __shfl_down(Lf.z, s, 64);
    }
    F3 l2 = L * (1.0f / 64.0f), fms = Lf * (1.0f / 64.0f);
    return make_float4(l2.x / (1.0f - fms.x), l2.y / (1.0f - fms.y), l2.z / (1.0f - fms.z), 1.0f);
}

// One multiple-scattering texel by the 64 lanes of a wave (lane = sphere direction);
// the result is valid in lane 0.  Every lane of the wave must call it.  kCells: T is the
// transmittance LUT's cell-corner copy.
template <bool kCells = false>
__device__ inline float4 multiscattering_texel(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                               int ix, int iy, int w, int h, int lane) {
    const MsRay ray = ms_ray(A, ix, iy, w, h, lane);
    F3 L, Tr, Lf;
    scattered_luminance<true, kCells>(A, T, tw, th, nullptr, 0, 0, ray.s, ray.dist, A.ms_steps, 0.5f, L, Tr, Lf);
    return multiscattering_reduce<kCells>(A, T, tw, th, ray, L, Tr, Lf);
}

// The renderer's block as the kernels see it.
struct RenderDev {
    F3 sun, up, right, front, cam;
    float r;
    float sv_steps, ap_steps, ap_max_distance;
};
__device__ inline RenderDev make_render(const SkyhipAtmosphereRenderBufferData& b) {
    RenderDev d;
    d.sun = f3(b.sun_direction);
    d.up = f3(b.up_direction);
    d.right = f3(b.right_direction);
    d.front = f3(b.front_direction);
    d.cam = f3(b.camera_position);
    d.r = b.camera_earth_center_distance;
    d.sv_steps = b.sky_view_lut_steps;
    d.ap_steps = b.aerial_perspective_lut_steps;
    d.ap_max_distance = b.aerial_perspective_lut_max_distance;
    return d;
}

// ComputeRaymarchingStartPositionAndChangeDistance (:56-72) on ray geometry:
// a camera above the top boundary starts at the entry point.
__device__ inline void start_from_space(const AtmDev& A, RayStart& s, float& marching_distance) {
    if (!(s.r > A.rt)) return;
    float disc = s.r * s.r * (s.mu * s.mu - 1.0f) + A.rt * A.rt;
    if (s.mu < 0.0f && disc >= 0.0f) {
        float near = fmaxf(-s.r * s.mu - sqrtf(fmaxf(disc, 0.0f)), 0.0f);
        float r2 = sqrtf(near * near + 2.0f * s.r * s.mu * near + s.r * s.r);
        s.mu = (s.r * s.mu + near) / r2;
        s.mu_s0 = (s.r * s.mu_s0 + near * s.nu) / r2;
        s.r = r2;
        s.alt0 = r2 - A.rb;
        marching_distance -= near;
    } else {
        marching_distance = 0.0f;
    }
}

// The ray of one sky-view texel (:74-186): its start, marching distance (0: nothing to
// integrate) and the float counter's start value.
struct SkyViewRay {
    RayStart s;
    float dist, start_i;
};
__device__ inline SkyViewRay sky_view_ray(const AtmDev& A, const RenderDev& R, const uint16_t* __restrict__ blue_noise,
                                          int bn_size, int ix, int iy, int w, int h) {
    const float r = R.r;
    const float alt0 = r - A.rb;
    // GetHorizonDownAngleFromR (:74-78) and GetCosLatLonFromSkyViewTextureIndex (:81-102).
    // The latitude is horizon_up + delta; delta is kept apart and cos/sin(lat) are formed
    // by angle addition from the horizon's exact cosine -rho/r and sine R_b/r, so rays a
    // few 1e-5 rad from the horizon (whose ground distance depends on delta steeply) keep
    // delta's full relative precision instead of an FP32 ulp of lat.
    const float rho = sqrtf(alt0 * (r + A.rb));
    const float horizon_down = acosf(rho / r);
    const float horizon_up = kPiF - horizon_down;
    const float cos_hu = -rho / r, sin_hu = A.rb / r;
    const float x_lon = float(ix) / float(w - 1);
    const float x_lat = float(iy) / float(h - 1);
    float delta;
    if (x_lat < 0.5f) {
        float c = 1.0f - 2.0f * x_lat;
        delta = -horizon_up * (c * c);  // lat = horizon_up (1 - c^2)
    } else {
        float c = x_lat * 2.0f - 1.0f;
        delta = horizon_down * (c * c);
    }
    float sin_d, cos_d;
    sincosf(delta, &sin_d, &cos_d);
    const float cos_lat = cos_hu * cos_d - sin_hu * sin_d;
    const float sin_lat = clamp01(sin_hu * cos_d + cos_hu * sin_d);
    const float cos_lon = -(x_lon * x_lon * 2.0f - 1.0f);
    const float sin_lon = clamp01(2.0f * x_lon * sqrtf(fmaxf(1.0f - x_lon * x_lon, 0.0f)));
    // GetViewDirectionFromCosLatLon (:104-111)
    const F3 view = R.up * cos_lat + R.front * (sin_lat * cos_lon) + R.right * (sin_lat * sin_lon);
    SkyViewRay ray{RayStart{r, alt0, cos_lat, dot(R.sun, R.up), dot(view, R.sun)}, 0.0f, 0.5f};
    const bool hit = ray_hits_ground(A, r, alt0, ray.s.mu);
    ray.dist = hit ? distance_to_bottom(A, r, alt0, ray.s.mu) : distance_to_top(A, r, alt0, ray.s.mu);
    start_from_space(A, ray.s, ray.dist);
    if (blue_noise) ray.start_i = float(blue_noise[(iy & (bn_size - 1)) * bn_size + (ix & (bn_size - 1))]) / 65535.0f;
    return ray;
}

// One sky-view texel (:74-186): luminance and transmittance (kCells: T and MS are cell-corner
// copies).
template <bool kCells = false>
__device__ inline void sky_view_texel(const AtmDev& A, const RenderDev& R, const float4* __restrict__ T, int tw,
                                      int th, const float4* __restrict__ MS, int mw, int mh,
                                      const uint16_t* __restrict__ blue_noise, int bn_size, int ix, int iy, int w,
                                      int h, F3& L, F3& Tr) {
    const SkyViewRay ray = sky_view_ray(A, R, blue_noise, bn_size, ix, iy, w, h);
    L = f3(0.0f);
    Tr = f3(1.0f);
    F3 Lf;
    if (ray.dist > 0.0f)
        scattered_luminance<false, kCells>(A, T, tw, th, MS, mw, mh, ray.s, ray.dist, R.sv_steps, ray.start_i, L, Tr,
                                           Lf);
}

// ProjectiveMul(inv_view_projection, (x, y, z)) - camera, normalised; the
// unprojection is done in FP64 (once per texel) so the direction keeps full
// FP32 accuracy however far the camera is from the world origin.
__device__ inline F3 unproject_direction(const float* m, float x, float y, float z, F3 cam, double* out_w_dist) {
    double v[4] = {x, y, z, 1.0};
    double o[4];
    for (int row = 0; row < 4; ++row)
        o[row] = double(m[row]) * v[0] + double(m[4 + row]) * v[1] + double(m[8 + row]) * v[2] +
                 double(m[12 + row]) * v[3];
    double px = o[0] / o[3] - double(cam.x), py = o[1] / o[3] - double(cam.y), pz = o[2] / o[3] - double(cam.z);
    double len = sqrt(px * px + py * py + pz * pz);
    if (out_w_dist) *out_w_dist = len;
    return {float(px / len), float(py / len), float(pz / len)};
}

// The set-up of one aerial-perspective froxel (:191-243): the ray start, the marching distance
// (0: nothing to integrate) and the float counter's start value.
struct AerialRay {
    RayStart s;
    float marching_distance, start_i;
};
__device__ inline AerialRay aerial_ray(const AtmDev& A, const RenderDev& R, const float* inv_view_projection,
                                       const uint16_t* __restrict__ blue_noise, int bn_size, int ix, int iy, int iz,
                                       int w, int h, int d) {
    const float u = float(ix) / float(w - 1), v = float(iy) / float(h - 1), q = float(iz) / float(d - 1);
    const F3 view = unproject_direction(inv_view_projection, u * 2.0f - 1.0f, v * 2.0f - 1.0f, 0.0f, R.cam, nullptr);
    float marching_distance = q * q * R.ap_max_distance;
    const float r = R.r;
    const float alt0 = r - A.rb;
    AerialRay a{RayStart{r, alt0, dot(view, R.up), dot(R.sun, R.up), dot(view, R.sun)}, 0.0f, 0.5f};
    const bool hit = ray_hits_ground(A, r, alt0, a.s.mu);
    float max_marching = hit ? marching_distance : distance_to_top(A, r, alt0, a.s.mu);
    start_from_space(A, a.s, max_marching);
    a.marching_distance = fminf(marching_distance, max_marching);
    if (blue_noise) a.start_i = float(blue_noise[(iy & (bn_size - 1)) * bn_size + (ix & (bn_size - 1))]) / 65535.0f;
    return a;
}

}  // namespace skyhip
