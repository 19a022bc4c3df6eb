// Test infrastructure (the oracle): C entry points over the templated CPU
// restatement, loaded with ctypes by tests/ and bench.py's cpu_baseline leg.
// *_f64 functions are the parity judge (FP64 arithmetic, FP64 outputs);
// *_f32 functions are the CPU baseline (the same source in FP32, OpenMP over
// texels / pixel rows, FP32 outputs).
#include <omp.h>

#include <vector>

#include "oracle_atmosphere.h"

using namespace oracle;

namespace {

template <class S, class O>
void transmittance_lut(const SkyhipAtmosphereBufferData* b, int w, int h, O* out) {
    Atmosphere<S> A(*b);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < w * h; ++t) {
        V3<S> v = A.TransmittanceTexel(t % w, t / w, w, h);
        out[t * 4 + 0] = O(v.x);
        out[t * 4 + 1] = O(v.y);
        out[t * 4 + 2] = O(v.z);
        out[t * 4 + 3] = O(1);
    }
}

template <class S, class O>
void multiscattering_lut(const SkyhipAtmosphereBufferData* b, const float* T, int tw, int th, int w, int h, O* out) {
    Atmosphere<S> A(*b);
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < w * h; ++t) {
        V3<S> v = A.MultiscatteringTexel(T, tw, th, t % w, t / w, w, h);
        out[t * 4 + 0] = O(v.x);
        out[t * 4 + 1] = O(v.y);
        out[t * 4 + 2] = O(v.z);
        out[t * 4 + 3] = O(1);
    }
}

template <class S, class O>
void sky_view_lut(const SkyhipAtmosphereBufferData* b, const SkyhipAtmosphereRenderBufferData* r, const float* T,
                  int tw, int th, const float* MS, int mw, int mh, const uint16_t* bn, int bn_size, int w, int h,
                  O* out_l, O* out_t) {
    Atmosphere<S> A(*b);
    RenderBlock<S> R(*r);
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < w * h; ++t) {
        V3<S> L, Tr;
        SkyViewTexel(A, R, T, tw, th, MS, mw, mh, bn, bn_size, t % w, t / w, w, h, L, Tr);
        out_l[t * 4 + 0] = O(L.x), out_l[t * 4 + 1] = O(L.y), out_l[t * 4 + 2] = O(L.z), out_l[t * 4 + 3] = O(0);
        out_t[t * 4 + 0] = O(Tr.x), out_t[t * 4 + 1] = O(Tr.y), out_t[t * 4 + 2] = O(Tr.z), out_t[t * 4 + 3] = O(0);
    }
}

template <class S, class O>
void aerial_perspective_lut(const SkyhipAtmosphereBufferData* b, const SkyhipAtmosphereRenderBufferData* r,
                            const float* T, int tw, int th, const float* MS, int mw, int mh, const uint16_t* bn,
                            int bn_size, int w, int h, int d, O* out_l, O* out_t) {
    Atmosphere<S> A(*b);
    RenderBlock<S> R(*r);
#pragma omp parallel for schedule(dynamic, 16)
    for (long t = 0; t < long(w) * h * d; ++t) {
        V3<S> L, Tr;
        int ix = int(t % w), iy = int((t / w) % h), iz = int(t / (long(w) * h));
        AerialPerspectiveTexel(A, R, T, tw, th, MS, mw, mh, bn, bn_size, ix, iy, iz, w, h, d, L, Tr);
        out_l[t * 4 + 0] = O(L.x), out_l[t * 4 + 1] = O(L.y), out_l[t * 4 + 2] = O(L.z), out_l[t * 4 + 3] = O(0);
        out_t[t * 4 + 0] = O(Tr.x), out_t[t * 4 + 1] = O(Tr.y), out_t[t * 4 + 2] = O(Tr.z), out_t[t * 4 + 3] = O(0);
    }
}

}  // namespace

extern "C" {

int oracle_max_threads() { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }

void oracle_transmittance_lut_f64(const SkyhipAtmosphereBufferData* b, int w, int h, double* out) {
    transmittance_lut<double>(b, w, h, out);
}
void oracle_transmittance_lut_f32(const SkyhipAtmosphereBufferData* b, int w, int h, float* out) {
    transmittance_lut<float>(b, w, h, out);
}
void oracle_multiscattering_lut_f64(const SkyhipAtmosphereBufferData* b, const float* T, int tw, int th, int w, int h,
                                    double* out) {
    multiscattering_lut<double>(b, T, tw, th, w, h, out);
}
void oracle_multiscattering_lut_f32(const SkyhipAtmosphereBufferData* b, const float* T, int tw, int th, int w, int h,
                                    float* out) {
    multiscattering_lut<float>(b, T, tw, th, w, h, out);
}
void oracle_sky_view_lut_f64(const SkyhipAtmosphereBufferData* b, const SkyhipAtmosphereRenderBufferData* r,
                             const float* T, int tw, int th, const float* MS, int mw, int mh, const uint16_t* bn,
                             int bn_size, int w, int h, double* out_l, double* out_t) {
    sky_view_lut<double>(b, r, T, tw, th, MS, mw, mh, bn, bn_size, w, h, out_l, out_t);
}
void oracle_sky_view_lut_f32(const SkyhipAtmosphereBufferData* b, const SkyhipAtmosphereRenderBufferData* r,
                             const float* T, int tw, int th, const float* MS, int mw, int mh, const uint16_t* bn,
                             int bn_size, int w, int h, float* out_l, float* out_t) {
    sky_view_lut<float>(b, r, T, tw, th, MS, mw, mh, bn, bn_size, w, h, out_l, out_t);
}
void oracle_aerial_perspective_lut_f64(const SkyhipAtmosphereBufferData* b, const SkyhipAtmosphereRenderBufferData* r,
                                       const float* T, int tw, int th, const float* MS, int mw, int mh,
                                       const uint16_t* bn, int bn_size, int w, int h, int d, double* out_l,
                                       double* out_t) {
    aerial_perspective_lut<double>(b, r, T, tw, th, MS, mw, mh, bn, bn_size, w, h, d, out_l, out_t);
}
void oracle_aerial_perspective_lut_f32(const SkyhipAtmosphereBufferData* b, const SkyhipAtmosphereRenderBufferData* r,
                                       const float* T, int tw, int th, const float* MS, int mw, int mh,
                                       const uint16_t* bn, int bn_size, int w, int h, int d, float* out_l,
                                       float* out_t) {
    aerial_perspective_lut<float>(b, r, T, tw, th, MS, mw, mh, bn, bn_size, w, h, d, out_l, out_t);
}

// Scalar probes used by the closed-form tests.
double oracle_extinction_f64(const SkyhipAtmosphereBufferData* b, double altitude, int channel) {
    Atmosphere<double> A(*b);
    return A.GetExtinction(altitude)[channel];
}
double oracle_mie_phase_f64(double g, double c) { return Atmosphere<double>::MiePhaseFunction(g, c); }
double oracle_rayleigh_phase_f64(double c) { return Atmosphere<double>::RayleighPhaseFunction(c); }
void oracle_transmittance_rmu_f64(const SkyhipAtmosphereBufferData* b, double r, double mu, double* out3) {
    Atmosphere<double> A(*b);
    V3<double> t = A.ComputeTransmittanceToTopAtmosphereBoundary(r, mu);
    out3[0] = t.x, out3[1] = t.y, out3[2] = t.z;
}
void oracle_transmittance_index_to_rmu_f64(const SkyhipAtmosphereBufferData* b, int ix, int iy, int w, int h,
                                           double* r, double* mu) {
    Atmosphere<double> A(*b);
    A.GetRMuFromTransmittanceTextureIndex(ix, iy, w, h, *r, *mu);
}
void oracle_transmittance_rmu_to_uv_f64(const SkyhipAtmosphereBufferData* b, int w, int h, double r, double mu,
                                        double* u, double* v) {
    Atmosphere<double> A(*b);
    A.GetTransmittanceTextureUvFromRMu(w, h, r, mu, *u, *v);
}
// Sky-view parameterisation round trip: texel index -> (cos_lat, cos_lon) -> uv.
void oracle_sky_view_index_roundtrip_f64(const SkyhipAtmosphereBufferData* b, double r, int ix, int iy, int w, int h,
                                         double* u, double* v) {
    Atmosphere<double> A(*b);
    double cl, co;
    GetCosLatLonFromSkyViewTextureIndex(A, ix, iy, w, h, r, cl, co);
    GetSkyViewTextureUvFromCosLatLon(A, w, h, r, cl, co, *u, *v);
}
// ComputeScatteredLuminance for a single ray (non-multiscattering variant).
void oracle_scattered_luminance_f64(const SkyhipAtmosphereBufferData* b, const float* T, int tw, int th,
                                    const float* MS, int mw, int mh, const double* start, const double* dir,
                                    const double* sun, double marching_distance, float steps, float start_i,
                                    double* out_l, double* out_t) {
    Atmosphere<double> A(*b);
    V3<double> c = {0.0, -double(b->bottom_radius), 0.0};
    V3<double> tr;
    V3<double> L = A.ComputeScatteredLuminance(false, T, tw, th, MS, mw, mh, start_i, c, {start[0], start[1], start[2]},
                                               {dir[0], dir[1], dir[2]}, {sun[0], sun[1], sun[2]}, marching_distance,
                                               steps, tr, nullptr);
    out_l[0] = L.x, out_l[1] = L.y, out_l[2] = L.z;
    out_t[0] = tr.x, out_t[1] = tr.y, out_t[2] = tr.z;
}

}  // extern "C"
