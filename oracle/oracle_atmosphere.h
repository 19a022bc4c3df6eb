// Test infrastructure (the oracle): CPU restatement of the atmosphere LUT
// math, shaders/SkyRendering/Atmosphere.glsl, AtmosphereRenderer.glsl:56-245
// and AtmosphereInterface.glsl, one function per GLSL function, in the
// reference's own (uncompensated) arithmetic.  Scalar type S = double is the
// parity judge, S = float the CPU baseline.
#pragma once

#include "../include/skyhip.h"
#include "oracle_common.h"

namespace oracle {

// The std140 AtmosphereBufferData block (Atmosphere.glsl:6-32) as scalars.
template <class S>
struct Atmosphere {
    V3<S> solar_illuminance;
    S sun_angular_radius;
    V3<S> rayleigh_scattering;
    S inv_rayleigh_exponential_distribution;
    V3<S> mie_scattering;
    S inv_mie_exponential_distribution;
    V3<S> mie_absorption;
    S ozone_center_altitude;
    V3<S> ozone_absorption;
    S inv_ozone_width;
    V3<S> ground_albedo;
    S mie_phase_g;
    S multiscattering_mask;
    S bottom_radius, top_radius, transmittance_steps, multiscattering_steps;

    explicit Atmosphere(const SkyhipAtmosphereBufferData& b)
        : solar_illuminance(cvt3<S>(b.solar_illuminance)),
          sun_angular_radius(b.sun_angular_radius),
          rayleigh_scattering(cvt3<S>(b.rayleigh_scattering)),
          inv_rayleigh_exponential_distribution(b.inv_rayleigh_exponential_distribution),
          mie_scattering(cvt3<S>(b.mie_scattering)),
          inv_mie_exponential_distribution(b.inv_mie_exponential_distribution),
          mie_absorption(cvt3<S>(b.mie_absorption)),
          ozone_center_altitude(b.ozone_center_altitude),
          ozone_absorption(cvt3<S>(b.ozone_absorption)),
          inv_ozone_width(b.inv_ozone_width),
          ground_albedo(cvt3<S>(b.ground_albedo)),
          mie_phase_g(b.mie_phase_g),
          multiscattering_mask(b.multiscattering_mask),
          bottom_radius(b.bottom_radius),
          top_radius(b.top_radius),
          transmittance_steps(b.transmittance_steps),
          multiscattering_steps(b.multiscattering_steps) {}

    // Atmosphere.glsl:37-51
    static S ClampCosine(S mu) { return gclamp(mu, S(-1), S(1)); }
    static S ClampDistance(S d) { return std::fmax(d, S(0)); }
    static S SafeSqrt(S a) { return std::sqrt(std::fmax(a, S(0))); }

    // Atmosphere.glsl:57-59
    bool RayIntersectsGround(S r, S mu) const {
        return mu < S(0) && r * r * (mu * mu - S(1)) + bottom_radius * bottom_radius >= S(0);
    }
    // Atmosphere.glsl:61-64
    S DistanceToTopAtmosphereBoundary(S r, S mu) const {
        S discriminant = r * r * (mu * mu - S(1)) + top_radius * top_radius;
        return ClampDistance(-r * mu + SafeSqrt(discriminant));
    }
    // Atmosphere.glsl:66-69
    S DistanceToBottomAtmosphereBoundary(S r, S mu) const {
        S discriminant = r * r * (mu * mu - S(1)) + bottom_radius * bottom_radius;
        return ClampDistance(-r * mu - SafeSqrt(discriminant));
    }

    // Atmosphere.glsl:71-88
    void GetRMuFromTransmittanceTextureIndex(int ix, int iy, int w, int h, S& r, S& mu) const {
        S x_mu = S(ix) / S(w - 1);
        S x_r = S(iy) / S(h - 1);
        S H = std::sqrt(top_radius * top_radius - bottom_radius * bottom_radius);
        S rho = H * x_r;
        r = std::sqrt(rho * rho + bottom_radius * bottom_radius);
        S d_min = top_radius - r;
        S d_max = rho + H;
        S d = d_min + x_mu * (d_max - d_min);
        mu = d == S(0) ? S(1) : (H * H - rho * rho - d * d) / (S(2) * r * d);
        mu = ClampCosine(mu);
    }

    // Atmosphere.glsl:90-103 (returns uv; GetTextureCoordFromUnitRange :53-55)
    void GetTransmittanceTextureUvFromRMu(int w, int h, S r, S mu, S& u, S& v) const {
        S H = std::sqrt(top_radius * top_radius - bottom_radius * bottom_radius);
        S rho = SafeSqrt(r * r - bottom_radius * bottom_radius);
        S d = DistanceToTopAtmosphereBoundary(r, mu);
        S d_min = top_radius - r;
        S d_max = rho + H;
        S x_mu = (d - d_min) / (d_max - d_min);
        S x_r = rho / H;
        u = S(0.5) / S(w) + x_mu * (S(1) - S(1) / S(w));
        v = S(0.5) / S(h) + x_r * (S(1) - S(1) / S(h));
    }

    // Atmosphere.glsl:105-117
    V3<S> GetSunVisibility(const float* T, int tw, int th, S r, S mu_s) const {
        S sin_theta_h = bottom_radius / r;
        S cos_theta_h = -std::sqrt(std::fmax(S(1) - sin_theta_h * sin_theta_h, S(0)));
        S u, v;
        GetTransmittanceTextureUvFromRMu(tw, th, r, mu_s, u, v);
        return tex2d_linear_clamp<S>(T, tw, th, u, v) *
               smoothstep(-sin_theta_h * sun_angular_radius, sin_theta_h * sun_angular_radius, mu_s - cos_theta_h);
    }

    // Atmosphere.glsl:119-132
    V3<S> GetExtinction(S altitude) const {
        V3<S> rayleigh_extinction =
            rayleigh_scattering * gclamp(std::exp(-altitude * inv_rayleigh_exponential_distribution), S(0), S(1));
        V3<S> mie_extinction = (mie_scattering + mie_absorption) *
                               gclamp(std::exp(-altitude * inv_mie_exponential_distribution), S(0), S(1));
        V3<S> ozone_extinction =
            ozone_absorption * std::fmax(S(0), altitude < ozone_center_altitude
                                                   ? S(1) + (altitude - ozone_center_altitude) * inv_ozone_width
                                                   : S(1) - (altitude - ozone_center_altitude) * inv_ozone_width);
        return rayleigh_extinction + mie_extinction + ozone_extinction;
    }

    // Atmosphere.glsl:134-154
    static S IsotropicPhaseFunction() { return S(1) / (S(4) * S(kPi)); }
    static S RayleighPhaseFunction(S cos_theta) {
        S k = S(3) / (S(16) * S(kPi));
        return k * (S(1) + cos_theta * cos_theta);
    }
    static S MiePhaseFunction(S g, S cos_theta) {
        S k = S(3) / (S(8) * S(kPi)) * (S(1) - g * g) / (S(2) + g * g);
        return k * (S(1) + cos_theta * cos_theta) / std::pow(S(1) + g * g - S(2) * g * cos_theta, S(1.5));
    }

    // Atmosphere.glsl:156-159
    void GetScattering(S altitude, V3<S>& rayleigh, V3<S>& mie) const {
        rayleigh = rayleigh_scattering * gclamp(std::exp(-altitude * inv_rayleigh_exponential_distribution), S(0), S(1));
        mie = mie_scattering * gclamp(std::exp(-altitude * inv_mie_exponential_distribution), S(0), S(1));
    }

    // Atmosphere.glsl:161-167
    void GetAltitudeMuSFromMultiscatteringTextureIndex(int ix, int iy, int w, int h, S& altitude, S& mu_s) const {
        S x_mu_s = S(ix) / S(w - 1);
        S x_altitude = S(iy) / S(h - 1);
        altitude = x_altitude * (top_radius - bottom_radius);
        mu_s = x_mu_s * S(2) - S(1);
    }

    // Atmosphere.glsl:169-178
    V3<S> GetMultiscatteringContribution(const float* MS, int mw, int mh, S r, S mu_s) const {
        S x_mu_s = mu_s * S(0.5) + S(0.5);
        S x_r = (r - bottom_radius) / (top_radius - bottom_radius);
        S u = S(0.5) / S(mw) + x_mu_s * (S(1) - S(1) / S(mw));
        S v = S(0.5) / S(mh) + x_r * (S(1) - S(1) / S(mh));
        return tex2d_linear_clamp<S>(MS, mw, mh, u, v);
    }

    // Atmosphere.glsl:220-295.  `multiscattering` selects the
    // MULTISCATTERING_COMPUTE_PROGRAM variant (isotropic phases, no solar
    // illuminance, no MS term, L_f accumulated, start_i = 0.5).  The float loop
    // counter of :255 is kept as a float.
    V3<S> ComputeScatteredLuminance(bool multiscattering, const float* T, int tw, int th, const float* MS, int mw,
                                    int mh, float start_i, V3<S> earth_center, V3<S> start_position,
                                    V3<S> view_direction, V3<S> sun_direction, S marching_distance, float steps,
                                    V3<S>& transmittance, V3<S>* L_f) const {
        S r = length(start_position - earth_center);
        V3<S> up_direction = normalize(start_position - earth_center);
        S mu = dot(view_direction, up_direction);
        S cos_sun_view = dot(view_direction, sun_direction);

        const S SAMPLE_COUNT = S(steps);
        S dx = marching_distance / SAMPLE_COUNT;

        transmittance = v3(S(1));
        V3<S> luminance = v3(S(0));
        if (multiscattering) {
            *L_f = v3(S(0));
            start_i = 0.5f;
        }
        S rayleigh_phase = multiscattering ? IsotropicPhaseFunction() : RayleighPhaseFunction(cos_sun_view);
        S mie_phase = multiscattering ? IsotropicPhaseFunction() : MiePhaseFunction(mie_phase_g, cos_sun_view);
        for (float i = start_i; i < steps; ++i) {
            S d_i = S(i) * dx;
            S r_i = std::sqrt(d_i * d_i + S(2) * r * mu * d_i + r * r);
            V3<S> position_i = start_position + view_direction * d_i;
            S altitude_i = r_i - bottom_radius;

            V3<S> rayleigh_scattering_i, mie_scattering_i;
            GetScattering(altitude_i, rayleigh_scattering_i, mie_scattering_i);
            V3<S> scattering_i = rayleigh_scattering_i + mie_scattering_i;
            V3<S> scattering_with_phase_i = rayleigh_scattering_i * rayleigh_phase + mie_scattering_i * mie_phase;

            V3<S> extinction_i = GetExtinction(altitude_i);
            V3<S> transmittance_i = vexp(-extinction_i * dx);
            V3<S> up_direction_i = normalize(position_i - earth_center);
            S mu_s_i = dot(sun_direction, up_direction_i);
            V3<S> luminance_i = scattering_with_phase_i * GetSunVisibility(T, tw, th, r_i, mu_s_i);
            if (!multiscattering) {
                V3<S> multiscattering_contribution = GetMultiscatteringContribution(MS, mw, mh, r_i, mu_s_i);
                luminance_i += multiscattering_mask * multiscattering_contribution * scattering_i;
                luminance_i *= solar_illuminance;
            }
            luminance += transmittance * (luminance_i - luminance_i * transmittance_i) / extinction_i;
            if (multiscattering)
                *L_f += transmittance * (scattering_i - scattering_i * transmittance_i) / extinction_i;
            transmittance *= transmittance_i;
        }
        return luminance;
    }

    // Atmosphere.glsl:297-306 (MS variant drops the solar illuminance).
    V3<S> ComputeGroundLuminance(bool multiscattering, const float* T, int tw, int th, V3<S> earth_center,
                                 V3<S> position, V3<S> sun_direction) const {
        V3<S> up_direction = normalize(position - earth_center);
        S mu_s = dot(sun_direction, up_direction);
        V3<S> solar_illuminance_at_ground = GetSunVisibility(T, tw, th, bottom_radius, mu_s);
        if (!multiscattering) solar_illuminance_at_ground *= solar_illuminance;
        V3<S> normal = normalize(position - earth_center);
        return S(1.0 / kPi) * gclamp(dot(normal, sun_direction), S(0), S(1)) * ground_albedo *
               solar_illuminance_at_ground;
    }

    // Atmosphere.glsl:311-327
    V3<S> ComputeTransmittanceToTopAtmosphereBoundary(S r, S mu) const {
        const S SAMPLE_COUNT = S(transmittance_steps);
        S dx = DistanceToTopAtmosphereBoundary(r, mu) / SAMPLE_COUNT;
        V3<S> optical_length = v3(S(0));
        float steps = float(transmittance_steps);
        for (float i = 0.5f; i < steps; ++i) {
            S d_i = S(i) * dx;
            S r_i = std::sqrt(d_i * d_i + S(2) * r * mu * d_i + r * r);
            S altitude_i = r_i - bottom_radius;
            optical_length += GetExtinction(altitude_i) * dx;
        }
        return vexp(-optical_length);
    }

    // Atmosphere.glsl:332-337 (one texel)
    V3<S> TransmittanceTexel(int ix, int iy, int w, int h) const {
        S r, mu;
        GetRMuFromTransmittanceTextureIndex(ix, iy, w, h, r, mu);
        return ComputeTransmittanceToTopAtmosphereBoundary(r, mu);
    }

    // Atmosphere.glsl:344-355
    static V3<S> GetDirectionFromLocalIndex(int index) {
        S unit_theta = (S(0.5) + S(index / 8)) / S(8);
        S unit_phi = (S(0.5) + S(index % 8)) / S(8);
        S cos_theta = S(1) - S(2) * unit_theta;
        S sin_theta = std::sqrt(gclamp(S(1) - cos_theta * cos_theta, S(0), S(1)));
        S phi = S(2) * S(kPi) * unit_phi;
        return {std::cos(phi) * sin_theta, cos_theta, std::sin(phi) * sin_theta};
    }

    // Atmosphere.glsl:364-438 (one texel; the 64 invocations and the fixed
    // shared-memory tree reduction of :392-427 are serialised in that order).
    V3<S> MultiscatteringTexel(const float* T, int tw, int th, int ix, int iy, int w, int h) const {
        S altitude, mu_s;
        GetAltitudeMuSFromMultiscatteringTextureIndex(ix, iy, w, h, altitude, mu_s);
        V3<S> earth_center = {S(0), -bottom_radius, S(0)};
        V3<S> start_position = {S(0), altitude, S(0)};
        V3<S> sun_direction = {S(0), mu_s, std::sqrt(S(1) - mu_s * mu_s)};
        V3<S> L2[64], Lf[64];
        for (int local_index = 0; local_index < 64; ++local_index) {
            V3<S> view_direction = GetDirectionFromLocalIndex(local_index);
            S r = altitude + bottom_radius;
            S mu = view_direction.y;
            bool intersect_bottom = RayIntersectsGround(r, mu);
            S marching_distance =
                intersect_bottom ? DistanceToBottomAtmosphereBoundary(r, mu) : DistanceToTopAtmosphereBoundary(r, mu);
            V3<S> transmittance, L_f;
            V3<S> luminance = ComputeScatteredLuminance(true, T, tw, th, nullptr, 0, 0, 0.5f, earth_center,
                                                        start_position, view_direction, sun_direction,
                                                        marching_distance, float(multiscattering_steps),
                                                        transmittance, &L_f);
            if (intersect_bottom) {
                V3<S> ground_position = start_position + view_direction * marching_distance;
                luminance += transmittance *
                             ComputeGroundLuminance(true, T, tw, th, earth_center, ground_position, sun_direction);
            }
            L2[local_index] = luminance;
            Lf[local_index] = L_f;
        }
        for (int stride = 32; stride >= 1; stride /= 2)
            for (int i = 0; i < stride; ++i) {
                L2[i] += L2[i + stride];
                Lf[i] += Lf[i + stride];
            }
        V3<S> L_2nd_order = L2[0] / S(64);
        V3<S> f_ms = Lf[0] / S(64);
        V3<S> F_ms = v3(S(1)) / (v3(S(1)) - f_ms);
        return L_2nd_order * F_ms;
    }
};

// AtmosphereRenderBufferData (AtmosphereRenderer.glsl:25-50), the fields the
// sky-view and aerial-perspective kernels read.
template <class S>
struct RenderBlock {
    V3<S> sun_direction, earth_center, camera_position, up_direction, right_direction, front_direction;
    S camera_earth_center_distance, sky_view_lut_steps_f, aerial_perspective_lut_steps_f,
        aerial_perspective_lut_max_distance;
    float sky_view_lut_steps, aerial_perspective_lut_steps;
    const float* inv_view_projection;

    explicit RenderBlock(const SkyhipAtmosphereRenderBufferData& b)
        : sun_direction(cvt3<S>(b.sun_direction)),
          earth_center(cvt3<S>(b.earth_center)),
          camera_position(cvt3<S>(b.camera_position)),
          up_direction(cvt3<S>(b.up_direction)),
          right_direction(cvt3<S>(b.right_direction)),
          front_direction(cvt3<S>(b.front_direction)),
          camera_earth_center_distance(b.camera_earth_center_distance),
          sky_view_lut_steps_f(b.sky_view_lut_steps),
          aerial_perspective_lut_steps_f(b.aerial_perspective_lut_steps),
          aerial_perspective_lut_max_distance(b.aerial_perspective_lut_max_distance),
          sky_view_lut_steps(b.sky_view_lut_steps),
          aerial_perspective_lut_steps(b.aerial_perspective_lut_steps),
          inv_view_projection(b.inv_view_projection) {}
};

// AtmosphereInterface.glsl:6-13
template <class S>
bool FromSpaceIntersectTopAtmosphereBoundary(const Atmosphere<S>& A, S r, S mu, S& near_distance) {
    S discriminant = r * r * (mu * mu - S(1)) + A.top_radius * A.top_radius;
    if (mu < S(0) && discriminant >= S(0)) {
        near_distance = Atmosphere<S>::ClampDistance(-r * mu - Atmosphere<S>::SafeSqrt(discriminant));
        return true;
    }
    return false;
}

// AtmosphereRenderer.glsl:56-72
template <class S>
V3<S> ComputeRaymarchingStartPositionAndChangeDistance(const Atmosphere<S>& A, const RenderBlock<S>& R,
                                                       V3<S> view_direction, S& marching_distance) {
    V3<S> start_position = R.camera_position;
    bool in_space = R.camera_earth_center_distance > A.top_radius;
    if (in_space) {
        S r = R.camera_earth_center_distance;
        S mu = dot(view_direction, R.up_direction);
        S near_distance;
        if (FromSpaceIntersectTopAtmosphereBoundary(A, r, mu, near_distance)) {
            start_position += near_distance * view_direction;
            marching_distance -= near_distance;
        } else {
            marching_distance = 0;
        }
    }
    return start_position;
}

// AtmosphereRenderer.glsl:74-78
template <class S>
S GetHorizonDownAngleFromR(const Atmosphere<S>& A, S r) {
    S tangent_point_distance = std::sqrt(r * r - A.bottom_radius * A.bottom_radius);
    S cos_horizon_down = tangent_point_distance / r;
    return std::acos(cos_horizon_down);
}

// AtmosphereRenderer.glsl:81-102 (SKY_VIEW_LUT_SIZE = (w, h))
template <class S>
void GetCosLatLonFromSkyViewTextureIndex(const Atmosphere<S>& A, int ix, int iy, int w, int h, S r, S& cos_lat,
                                         S& cos_lon) {
    S x_cos_lon = S(ix) / S(w - 1);
    S x_cos_lat = S(iy) / S(h - 1);
    S horizon_down_angle = GetHorizonDownAngleFromR(A, r);
    S horizon_up_angle = S(kPi) - horizon_down_angle;
    S lat;
    if (x_cos_lat < S(0.5)) {
        S coord = S(1) - S(2) * x_cos_lat;
        coord = S(1) - coord * coord;
        lat = horizon_up_angle * coord;
    } else {
        S coord = x_cos_lat * S(2) - S(1);
        coord *= coord;
        lat = horizon_up_angle + horizon_down_angle * coord;
    }
    cos_lat = std::cos(lat);
    cos_lon = -(x_cos_lon * x_cos_lon * S(2) - S(1));
}

// AtmosphereRenderer.glsl:104-111
template <class S>
V3<S> GetViewDirectionFromCosLatLon(const RenderBlock<S>& R, S cos_lat, S cos_lon) {
    S sin_lat = gclamp(std::sqrt(S(1) - cos_lat * cos_lat), S(0), S(1));
    S sin_lon = gclamp(std::sqrt(S(1) - cos_lon * cos_lon), S(0), S(1));
    return R.up_direction * cos_lat + R.front_direction * (sin_lat * cos_lon) +
           R.right_direction * (sin_lat * sin_lon);
}

// AtmosphereRenderer.glsl:113-132 (uv of a (cos_lat, cos_lon) pair in a w x h LUT)
template <class S>
void GetSkyViewTextureUvFromCosLatLon(const Atmosphere<S>& A, int w, int h, S r, S cos_lat, S cos_lon, S& u, S& v) {
    S horizon_down_angle = GetHorizonDownAngleFromR(A, r);
    S horizon_up_angle = S(kPi) - horizon_down_angle;
    S lat = std::acos(cos_lat);
    S x_cos_lat;
    if (lat < horizon_up_angle) {
        S coord = lat / horizon_up_angle;
        coord = std::sqrt(S(1) - coord);
        x_cos_lat = S(0.5) - S(0.5) * coord;
    } else {
        S coord = (lat - horizon_up_angle) / horizon_down_angle;
        coord = std::sqrt(coord);
        x_cos_lat = coord * S(0.5) + S(0.5);
    }
    S x_cos_lon = std::sqrt(S(0.5) - S(0.5) * cos_lon);
    u = S(0.5) / S(w) + x_cos_lon * (S(1) - S(1) / S(w));
    v = S(0.5) / S(h) + x_cos_lat * (S(1) - S(1) / S(h));
}

// AtmosphereRenderer.glsl:134-145
template <class S>
void GetCosLatLonFromViewDirection(const RenderBlock<S>& R, V3<S> view_direction, S& cos_lat, S& cos_lon) {
    cos_lat = dot(R.up_direction, view_direction);
    V3<S> lon_direction = view_direction - R.up_direction * cos_lat;
    S l2 = dot(lon_direction, lon_direction);
    if (l2 == S(0)) {
        cos_lon = S(1);
    } else {
        lon_direction *= S(1) / std::sqrt(l2);
        cos_lon = dot(lon_direction, R.front_direction);
    }
}

// Sky-view texel, AtmosphereRenderer.glsl:153-186.
template <class S>
void SkyViewTexel(const Atmosphere<S>& A, const RenderBlock<S>& R, const float* T, int tw, int th, const float* MS,
                  int mw, int mh, const uint16_t* blue_noise, int bn_size, int ix, int iy, int w, int h,
                  V3<S>& luminance, V3<S>& transmittance) {
    S r = R.camera_earth_center_distance;
    S cos_lat, cos_lon;
    GetCosLatLonFromSkyViewTextureIndex(A, ix, iy, w, h, r, cos_lat, cos_lon);
    V3<S> view_direction = GetViewDirectionFromCosLatLon(R, cos_lat, cos_lon);
    S mu = cos_lat;
    bool intersect_bottom = A.RayIntersectsGround(r, mu);
    S marching_distance =
        intersect_bottom ? A.DistanceToBottomAtmosphereBoundary(r, mu) : A.DistanceToTopAtmosphereBoundary(r, mu);
    V3<S> start_position = ComputeRaymarchingStartPositionAndChangeDistance(A, R, view_direction, marching_distance);
    transmittance = v3(S(1));
    luminance = v3(S(0));
    if (marching_distance > S(0)) {
        float start_i = 0.5f;
        if (blue_noise) start_i = float(blue_noise[(iy & (bn_size - 1)) * bn_size + (ix & (bn_size - 1))]) / 65535.0f;
        luminance = A.ComputeScatteredLuminance(false, T, tw, th, MS, mw, mh, start_i, R.earth_center, start_position,
                                                view_direction, R.sun_direction, marching_distance,
                                                R.sky_view_lut_steps, transmittance, nullptr);
    }
}

// Aerial-perspective texel, AtmosphereRenderer.glsl:191-243.
template <class S>
void AerialPerspectiveTexel(const Atmosphere<S>& A, const RenderBlock<S>& R, const float* T, int tw, int th,
                            const float* MS, int mw, int mh, const uint16_t* blue_noise, int bn_size, int ix, int iy,
                            int iz, int w, int h, int d, V3<S>& luminance, V3<S>& transmittance) {
    // GetMarchingDistanceFromAerialPerspectiveTextureIndex, :191-197
    V3<S> uvw = {S(ix) / S(w - 1), S(iy) / S(h - 1), S(iz) / S(d - 1)};
    V3<S> position = projective_mul<S>(R.inv_view_projection, {uvw.x * S(2) - S(1), uvw.y * S(2) - S(1), S(0)});
    V3<S> view_direction = normalize(position - R.camera_position);
    S marching_distance = uvw.z * uvw.z * R.aerial_perspective_lut_max_distance;

    S r = R.camera_earth_center_distance;
    S mu = dot(view_direction, R.up_direction);
    bool intersect_bottom = A.RayIntersectsGround(r, mu);
    S max_marching_distance = intersect_bottom ? marching_distance : A.DistanceToTopAtmosphereBoundary(r, mu);
    V3<S> start_position =
        ComputeRaymarchingStartPositionAndChangeDistance(A, R, view_direction, max_marching_distance);
    marching_distance = std::fmin(marching_distance, max_marching_distance);
    transmittance = v3(S(1));
    luminance = v3(S(0));
    if (marching_distance > S(0)) {
        float start_i = 0.5f;
        if (blue_noise) start_i = float(blue_noise[(iy & (bn_size - 1)) * bn_size + (ix & (bn_size - 1))]) / 65535.0f;
        luminance = A.ComputeScatteredLuminance(false, T, tw, th, MS, mw, mh, start_i, R.earth_center, start_position,
                                                view_direction, R.sun_direction, marching_distance,
                                                R.aerial_perspective_lut_steps, transmittance, nullptr);
    }
}

}  // namespace oracle
