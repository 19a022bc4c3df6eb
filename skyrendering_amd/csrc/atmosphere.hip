// Atmosphere LUT kernels for gfx950: transmittance, multiple scattering,
// sky view and aerial perspective, plus their C-ABI entry points.
//
// Layout: every LUT is RGBA32F in GL texel order, one lane per texel (one
// wave64 per texel for multiple scattering, whose 64 sphere directions map to
// the 64 lanes), 256-lane workgroups, 1D grids over texels so stores coalesce.
#include "atmosphere_device.h"

#include <cmath>

namespace skyhip {

// ---------------------------------------------------------------------------
// Transmittance (Atmosphere.glsl:71-88, :311-337)
// ---------------------------------------------------------------------------
__device__ inline float sexp(float x) { return __expf(x); }
__device__ inline double sexp(double x) { return exp(x); }
__device__ inline float ssqrt(float x) { return sqrtf(x); }
__device__ inline double ssqrt(double x) { return sqrt(x); }

// One texel, evaluated entirely in scalar type S.
template <class S>
__device__ inline F3 transmittance_texel(const SkyhipAtmosphereBufferData& b, int ix, int iy, int w, int h) {
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
    S od_r = 0, od_m = 0, od_o = 0;  // density integrals (extinction is linear in them)
    for (float i = 0.5f; i < steps; i += 1.0f) {
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
    F3 out;
    for (int c = 0; c < 3; ++c) {
        S od = (S(b.rayleigh_scattering[c]) * od_r + (S(b.mie_scattering[c]) + S(b.mie_absorption[c])) * od_m +
                S(b.ozone_absorption[c]) * od_o) * dx;
        (&out.x)[c] = float(sexp(-od));
    }
    return out;
}

// fp64_from_column: texels with x >= this column (the near-horizon end of the
// x_mu axis, where the optical depth is largest) integrate in FP64.
__global__ __launch_bounds__(256) void transmittance_kernel(SkyhipAtmosphereBufferData b, float4* __restrict__ out,
                                                            int w, int h, int fp64_from_column) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w * h) return;
    int ix = t % w, iy = t / w;
    F3 v = ix >= fp64_from_column ? transmittance_texel<double>(b, ix, iy, w, h)
                                  : transmittance_texel<float>(b, ix, iy, w, h);
    out[t] = make_float4(v.x, v.y, v.z, 1.0f);
}

// ---------------------------------------------------------------------------
// Multiple scattering (Atmosphere.glsl:344-438): one wave64 per texel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void multiscattering_kernel(SkyhipAtmosphereBufferData b,
                                                              const float4* __restrict__ T, int tw, int th,
                                                              float4* __restrict__ out, int w, int h) {
    const AtmDev A = make_atm(b);
    const int texel = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (texel >= w * h) return;  // wave-uniform
    const int ix = texel % w, iy = texel / w;
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
    F3 L, Tr, Lf;
    scattered_luminance<true>(A, T, tw, th, nullptr, 0, 0, RayStart{r, altitude, mu, mu_s, nu}, dist, A.ms_steps,
                              0.5f, L, Tr, Lf);
    if (hit) {
        // ComputeGroundLuminance (:297-306) at the ground point, |p - c| = R_b.
        float mu_g = (r * mu_s + dist * nu) / A.rb;
        F3 sv = sun_visibility(A, T, tw, th, A.rb, 0.0f, 0.0f, mu_g);
        L += Tr * ((1.0f / kPiF) * clamp01(mu_g)) * A.albedo * sv;
    }
    // The reference's fixed-order shared-memory tree (:392-427): lane i adds
    // lane i + s for s = 32, 16, ..., 1, which is exactly this shuffle chain.
    for (int s = 32; s >= 1; s >>= 1) {
        L.x += __shfl_down(L.x, s, 64);
        L.y += __shfl_down(L.y, s, 64);
        L.z += __shfl_down(L.z, s, 64);
        Lf.x += __shfl_down(Lf.x, s, 64);
        Lf.y += __shfl_down(Lf.y, s, 64);
        Lf.z += __shfl_down(Lf.z, s, 64);
    }
    if (lane == 0) {
        F3 l2 = L * (1.0f / 64.0f), fms = Lf * (1.0f / 64.0f);
        out[texel] = make_float4(l2.x / (1.0f - fms.x), l2.y / (1.0f - fms.y), l2.z / (1.0f - fms.z), 1.0f);
    }
}

// ---------------------------------------------------------------------------
// Sky view (AtmosphereRenderer.glsl:74-186)
// ---------------------------------------------------------------------------
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

__global__ __launch_bounds__(256) void sky_view_kernel(SkyhipAtmosphereBufferData b,
                                                       SkyhipAtmosphereRenderBufferData rbd,
                                                       const float4* __restrict__ T, int tw, int th,
                                                       const float4* __restrict__ MS, int mw, int mh,
                                                       const uint16_t* __restrict__ blue_noise, int bn_size,
                                                       float4* __restrict__ out_l, float4* __restrict__ out_t, int w,
                                                       int h) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w * h) return;
    const AtmDev A = make_atm(b);
    const RenderDev R = make_render(rbd);
    const int ix = t % w, iy = t / w;
    const float r = R.r;
    const float alt0 = r - A.rb;
    // GetHorizonDownAngleFromR (:74-78) and GetCosLatLonFromSkyViewTextureIndex (:81-102)
    const float horizon_down = acosf(sqrtf(alt0 * (r + A.rb)) / r);
    const float horizon_up = kPiF - horizon_down;
    const float x_lon = float(ix) / float(w - 1);
    const float x_lat = float(iy) / float(h - 1);
    float lat;
    if (x_lat < 0.5f) {
        float c = 1.0f - 2.0f * x_lat;
        lat = horizon_up * (1.0f - c * c);
    } else {
        float c = x_lat * 2.0f - 1.0f;
        lat = horizon_up + horizon_down * (c * c);
    }
    float sin_lat, cos_lat;
    sincosf(lat, &sin_lat, &cos_lat);
    sin_lat = clamp01(sin_lat);
    const float cos_lon = -(x_lon * x_lon * 2.0f - 1.0f);
    const float sin_lon = clamp01(2.0f * x_lon * sqrtf(fmaxf(1.0f - x_lon * x_lon, 0.0f)));
    // GetViewDirectionFromCosLatLon (:104-111)
    const F3 view = R.up * cos_lat + R.front * (sin_lat * cos_lon) + R.right * (sin_lat * sin_lon);
    RayStart s{r, alt0, cos_lat, dot(R.sun, R.up), dot(view, R.sun)};
    const bool hit = ray_hits_ground(A, r, alt0, s.mu);
    float dist = hit ? distance_to_bottom(A, r, alt0, s.mu) : distance_to_top(A, r, alt0, s.mu);
    start_from_space(A, s, dist);
    F3 L = f3(0.0f), Tr = f3(1.0f), Lf;
    if (dist > 0.0f) {
        float start_i = 0.5f;
        if (blue_noise) start_i = float(blue_noise[(iy & (bn_size - 1)) * bn_size + (ix & (bn_size - 1))]) / 65535.0f;
        scattered_luminance<false>(A, T, tw, th, MS, mw, mh, s, dist, R.sv_steps, start_i, L, Tr, Lf);
    }
    out_l[t] = make_float4(L.x, L.y, L.z, 0.0f);
    out_t[t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
}

// ---------------------------------------------------------------------------
// Aerial perspective (AtmosphereRenderer.glsl:191-243)
// ---------------------------------------------------------------------------
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

__global__ __launch_bounds__(256) void aerial_perspective_kernel(
    SkyhipAtmosphereBufferData b, SkyhipAtmosphereRenderBufferData rbd, const float4* __restrict__ T, int tw, int th,
    const float4* __restrict__ MS, int mw, int mh, const uint16_t* __restrict__ blue_noise, int bn_size,
    float4* __restrict__ out_l, float4* __restrict__ out_t, int w, int h, int d) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= w * h * d) return;
    const AtmDev A = make_atm(b);
    const RenderDev R = make_render(rbd);
    const int ix = t % w, iy = (t / w) % h, iz = t / (w * h);
    const float u = float(ix) / float(w - 1), v = float(iy) / float(h - 1), q = float(iz) / float(d - 1);
    const F3 view = unproject_direction(rbd.inv_view_projection, u * 2.0f - 1.0f, v * 2.0f - 1.0f, 0.0f, R.cam,
                                        nullptr);
    float marching_distance = q * q * R.ap_max_distance;
    const float r = R.r;
    const float alt0 = r - A.rb;
    RayStart s{r, alt0, dot(view, R.up), dot(R.sun, R.up), dot(view, R.sun)};
    const bool hit = ray_hits_ground(A, r, alt0, s.mu);
    float max_marching = hit ? marching_distance : distance_to_top(A, r, alt0, s.mu);
    start_from_space(A, s, max_marching);
    marching_distance = fminf(marching_distance, max_marching);
    F3 L = f3(0.0f), Tr = f3(1.0f), Lf;
    if (marching_distance > 0.0f) {
        float start_i = 0.5f;
        if (blue_noise) start_i = float(blue_noise[(iy & (bn_size - 1)) * bn_size + (ix & (bn_size - 1))]) / 65535.0f;
        scattered_luminance<false>(A, T, tw, th, MS, mw, mh, s, marching_distance, R.ap_steps, start_i, L, Tr, Lf);
    }
    out_l[t] = make_float4(L.x, L.y, L.z, 0.0f);
    out_t[t] = make_float4(Tr.x, Tr.y, Tr.z, 0.0f);
}

// ---------------------------------------------------------------------------
// Validation and launch helpers
// ---------------------------------------------------------------------------
constexpr int kMaxLutDim = 8192;

int validate_atmosphere(const SkyhipAtmosphereBufferData* b) {
    SKYHIP_CHECK(require(b != nullptr, SKYHIP_ERR_INVALID_ARGUMENT, "atmosphere block is NULL"));
    SKYHIP_CHECK(require(finite_block(reinterpret_cast<const float*>(b), sizeof(*b) / 4), SKYHIP_ERR_NON_FINITE,
                         "atmosphere block has a non-finite field"));
    SKYHIP_CHECK(require(b->bottom_radius > 0.0f && b->top_radius > b->bottom_radius, SKYHIP_ERR_NON_FINITE,
                         "atmosphere needs 0 < bottom_radius < top_radius"));
    SKYHIP_CHECK(require(b->transmittance_steps > 0.0f && b->transmittance_steps <= 4096.0f &&
                             b->multiscattering_steps > 0.0f && b->multiscattering_steps <= 4096.0f,
                         SKYHIP_ERR_INVALID_SIZE, "step counts must be in (0, 4096]"));
    return SKYHIP_OK;
}

int validate_render(const SkyhipAtmosphereRenderBufferData* r) {
    SKYHIP_CHECK(require(r != nullptr, SKYHIP_ERR_INVALID_ARGUMENT, "render block is NULL"));
    SKYHIP_CHECK(require(finite_block(reinterpret_cast<const float*>(r), sizeof(*r) / 4), SKYHIP_ERR_NON_FINITE,
                         "render block has a non-finite field"));
    SKYHIP_CHECK(require(r->sky_view_lut_steps >= 0.0f && r->sky_view_lut_steps <= 4096.0f &&
                             r->aerial_perspective_lut_steps >= 0.0f && r->aerial_perspective_lut_steps <= 4096.0f,
                         SKYHIP_ERR_INVALID_SIZE, "step counts must be in [0, 4096]"));
    return SKYHIP_OK;
}

int validate_lut(const void* p, int w, int h, const char* what) {
    if (p == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "%s pointer is NULL", what);
    if (w < 2 || h < 2 || w > kMaxLutDim || h > kMaxLutDim)
        return fail(SKYHIP_ERR_INVALID_SIZE, "%s size %dx%d outside [2, %d]", what, w, h, kMaxLutDim);
    return SKYHIP_OK;
}

int validate_blue_noise(const uint16_t* bn, int size) {
    if (bn == nullptr) return SKYHIP_OK;
    if (size < 1 || size > 4096 || (size & (size - 1)) != 0)
        return fail(SKYHIP_ERR_INVALID_SIZE, "blue-noise tile size %d is not a power of two in [1, 4096]", size);
    return SKYHIP_OK;
}

int launch_transmittance(hipStream_t st, const SkyhipAtmosphereBufferData& b, float* out, int w, int h) {
    hipLaunchKernelGGL(transmittance_kernel, dim3(div_up(w * h, 256)), dim3(256), 0, st, b,
                       reinterpret_cast<float4*>(out), w, h, (w * 3) / 4);
    SKYHIP_LAUNCH_CHECK("transmittance_kernel");
    return SKYHIP_OK;
}

int launch_multiscattering(hipStream_t st, const SkyhipAtmosphereBufferData& b, const float* T, int tw, int th,
                           float* out, int w, int h) {
    hipLaunchKernelGGL(multiscattering_kernel, dim3(div_up(w * h, 4)), dim3(256), 0, st, b,
                       reinterpret_cast<const float4*>(T), tw, th, reinterpret_cast<float4*>(out), w, h);
    SKYHIP_LAUNCH_CHECK("multiscattering_kernel");
    return SKYHIP_OK;
}

int launch_sky_view(hipStream_t st, const SkyhipAtmosphereBufferData& b, const SkyhipAtmosphereRenderBufferData& r,
                    const float* T, int tw, int th, const float* MS, int mw, int mh, const uint16_t* bn, int bn_size,
                    float* L, float* Tr, int w, int h) {
    hipLaunchKernelGGL(sky_view_kernel, dim3(div_up(w * h, 256)), dim3(256), 0, st, b, r,
                       reinterpret_cast<const float4*>(T), tw, th, reinterpret_cast<const float4*>(MS), mw, mh, bn,
                       bn_size, reinterpret_cast<float4*>(L), reinterpret_cast<float4*>(Tr), w, h);
    SKYHIP_LAUNCH_CHECK("sky_view_kernel");
    return SKYHIP_OK;
}

int launch_aerial_perspective(hipStream_t st, const SkyhipAtmosphereBufferData& b,
                              const SkyhipAtmosphereRenderBufferData& r, const float* T, int tw, int th,
                              const float* MS, int mw, int mh, const uint16_t* bn, int bn_size, float* L, float* Tr,
                              int w, int h, int d) {
    hipLaunchKernelGGL(aerial_perspective_kernel, dim3(div_up((long long)w * h * d, 256)), dim3(256), 0, st, b, r,
                       reinterpret_cast<const float4*>(T), tw, th, reinterpret_cast<const float4*>(MS), mw, mh, bn,
                       bn_size, reinterpret_cast<float4*>(L), reinterpret_cast<float4*>(Tr), w, h, d);
    SKYHIP_LAUNCH_CHECK("aerial_perspective_kernel");
    return SKYHIP_OK;
}

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_transmittance_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                        float* transmittance_rgba, int32_t width, int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, width, height, "transmittance"));
    return launch_transmittance(ctx->stream, *atmosphere, transmittance_rgba, width, height);
}

extern "C" int skyhip_multiscattering_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                          const float* transmittance_rgba, int32_t transmittance_width,
                                          int32_t transmittance_height, float* multiscattering_rgba, int32_t width,
                                          int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, width, height, "multiscattering"));
    return launch_multiscattering(ctx->stream, *atmosphere, transmittance_rgba, transmittance_width,
                                  transmittance_height, multiscattering_rgba, width, height);
}

extern "C" int skyhip_atmosphere_update_luts(SkyhipContext* ctx, const SkyhipAtmosphereParameters* parameters,
                                             float* transmittance_rgba, int32_t transmittance_width,
                                             int32_t transmittance_height, float* multiscattering_rgba,
                                             int32_t multiscattering_width, int32_t multiscattering_height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SkyhipAtmosphereBufferData b;
    SKYHIP_CHECK(skyhip_atmosphere_assign_buffer_data(parameters, &b));
    SKYHIP_CHECK(validate_atmosphere(&b));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, multiscattering_width, multiscattering_height, "multiscattering"));
    SKYHIP_CHECK(launch_transmittance(ctx->stream, b, transmittance_rgba, transmittance_width, transmittance_height));
    return launch_multiscattering(ctx->stream, b, transmittance_rgba, transmittance_width, transmittance_height,
                                  multiscattering_rgba, multiscattering_width, multiscattering_height);
}

extern "C" int skyhip_sky_view_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                   const SkyhipAtmosphereRenderBufferData* render, const float* transmittance_rgba,
                                   int32_t transmittance_width, int32_t transmittance_height,
                                   const float* multiscattering_rgba, int32_t multiscattering_width,
                                   int32_t multiscattering_height, const uint16_t* blue_noise, int32_t blue_noise_size,
                                   float* luminance_rgba, float* sky_transmittance_rgba, int32_t width,
                                   int32_t height) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, multiscattering_width, multiscattering_height, "multiscattering"));
    SKYHIP_CHECK(validate_blue_noise(blue_noise, blue_noise_size));
    SKYHIP_CHECK(validate_lut(luminance_rgba, width, height, "sky-view luminance"));
    SKYHIP_CHECK(validate_lut(sky_transmittance_rgba, width, height, "sky-view transmittance"));
    return launch_sky_view(ctx->stream, *atmosphere, *render, transmittance_rgba, transmittance_width,
                           transmittance_height, multiscattering_rgba, multiscattering_width, multiscattering_height,
                           blue_noise, blue_noise_size, luminance_rgba, sky_transmittance_rgba, width, height);
}

extern "C" int skyhip_aerial_perspective_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                             const SkyhipAtmosphereRenderBufferData* render,
                                             const float* transmittance_rgba, int32_t transmittance_width,
                                             int32_t transmittance_height, const float* multiscattering_rgba,
                                             int32_t multiscattering_width, int32_t multiscattering_height,
                                             const uint16_t* blue_noise, int32_t blue_noise_size,
                                             float* luminance_rgba, float* ap_transmittance_rgba, int32_t width,
                                             int32_t height, int32_t depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, multiscattering_width, multiscattering_height, "multiscattering"));
    SKYHIP_CHECK(validate_blue_noise(blue_noise, blue_noise_size));
    SKYHIP_CHECK(validate_lut(luminance_rgba, width, height, "aerial-perspective luminance"));
    SKYHIP_CHECK(validate_lut(ap_transmittance_rgba, width, height, "aerial-perspective transmittance"));
    if (depth < 2 || depth > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "aerial-perspective depth %d outside [2, 1024]", depth);
    return launch_aerial_perspective(ctx->stream, *atmosphere, *render, transmittance_rgba, transmittance_width,
                                     transmittance_height, multiscattering_rgba, multiscattering_width,
                                     multiscattering_height, blue_noise, blue_noise_size, luminance_rgba,
                                     ap_transmittance_rgba, width, height, depth);
}

extern "C" int skyhip_atmosphere_lut_set(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                         const SkyhipAtmosphereRenderBufferData* render, float* transmittance_rgba,
                                         int32_t transmittance_width, int32_t transmittance_height,
                                         float* multiscattering_rgba, int32_t multiscattering_width,
                                         int32_t multiscattering_height, float* sky_view_luminance_rgba,
                                         float* sky_view_transmittance_rgba, int32_t sky_view_width,
                                         int32_t sky_view_height, float* ap_luminance_rgba,
                                         float* ap_transmittance_rgba, int32_t ap_width, int32_t ap_height,
                                         int32_t ap_depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    SKYHIP_CHECK(validate_render(render));
    const int tw = transmittance_width, th = transmittance_height, mw = multiscattering_width,
              mh = multiscattering_height;
    SKYHIP_CHECK(validate_lut(transmittance_rgba, tw, th, "transmittance"));
    SKYHIP_CHECK(validate_lut(multiscattering_rgba, mw, mh, "multiscattering"));
    SKYHIP_CHECK(validate_lut(sky_view_luminance_rgba, sky_view_width, sky_view_height, "sky-view luminance"));
    SKYHIP_CHECK(validate_lut(sky_view_transmittance_rgba, sky_view_width, sky_view_height, "sky-view transmittance"));
    SKYHIP_CHECK(validate_lut(ap_luminance_rgba, ap_width, ap_height, "aerial-perspective luminance"));
    SKYHIP_CHECK(validate_lut(ap_transmittance_rgba, ap_width, ap_height, "aerial-perspective transmittance"));
    if (ap_depth < 2 || ap_depth > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "aerial-perspective depth outside [2, 1024]");
    hipStream_t st = ctx->stream;
    SKYHIP_CHECK(launch_transmittance(st, *atmosphere, transmittance_rgba, tw, th));
    SKYHIP_CHECK(launch_multiscattering(st, *atmosphere, transmittance_rgba, tw, th, multiscattering_rgba, mw, mh));
    // Sky view and aerial perspective both read T and MS only: run them side by side.
    SKYHIP_HIP(hipEventRecord(ctx->fork_event, st));
    SKYHIP_HIP(hipStreamWaitEvent(ctx->aux_stream, ctx->fork_event, 0));
    SKYHIP_CHECK(launch_aerial_perspective(ctx->aux_stream, *atmosphere, *render, transmittance_rgba, tw, th,
                                           multiscattering_rgba, mw, mh, nullptr, 1, ap_luminance_rgba,
                                           ap_transmittance_rgba, ap_width, ap_height, ap_depth));
    SKYHIP_CHECK(launch_sky_view(st, *atmosphere, *render, transmittance_rgba, tw, th, multiscattering_rgba, mw, mh,
                                 nullptr, 1, sky_view_luminance_rgba, sky_view_transmittance_rgba, sky_view_width,
                                 sky_view_height));
    SKYHIP_HIP(hipEventRecord(ctx->join_event, ctx->aux_stream));
    SKYHIP_HIP(hipStreamWaitEvent(st, ctx->join_event, 0));
    return SKYHIP_OK;
}
