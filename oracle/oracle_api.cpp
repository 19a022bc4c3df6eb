// Test infrastructure (the oracle): C entry points over the templated CPU
// restatement, loaded with ctypes by tests/ and bench.py's cpu_baseline leg.
// *_f64 functions are the parity judge (FP64 arithmetic, FP64 outputs);
// *_f32 functions are the CPU baseline (the same source in FP32, OpenMP over
// texels / pixel rows, FP32 outputs).
#include <omp.h>

#include <vector>

#include "oracle_cloud.h"
#include "oracle_noise.h"
#include "oracle_cloud_shadow.h"
#include "oracle_cloud_reconstruct.h"
#include "oracle_ibl.h"
#include "oracle_pathtrace.h"

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

template <class S, class O>
void cloud_raymarch(const SkyhipAtmosphereBufferData* atm, const SkyhipCloudCommonBufferData* common,
                           const SkyhipCloudBufferData* cloud, const SkyhipCloudMaterial* material,
                           const CloudImages* images, int width, const int32_t* pixels, long count, O* out_rgba,
                           O* out_dist, int32_t* out_steps, const uint32_t* forced_steps = nullptr,
                           double* out_margin = nullptr, double* out_step_margin = nullptr,
                           double* out_texel_margin = nullptr, double* out_lod_margin = nullptr) {
    Atmosphere<S> A(*atm);
    CloudCtx<S> Cc(common, cloud, material);
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < count; ++i) {
        int p = pixels[i];
        S margin = 0, step_margin = 0;
        DecisionMargins fetch_margins;
        if (out_texel_margin || out_lod_margin) ActiveMargins() = &fetch_margins;
        CloudPixelOut<S> o = CloudRaymarchPixel<S>(A, Cc, *images, p % width, p / width,
                                                   forced_steps ? forced_steps + 2 * i : nullptr, &margin, nullptr, 0,
                                                   &step_margin);
        ActiveMargins() = nullptr;
        if (out_texel_margin) out_texel_margin[i] = fetch_margins.texel;
        if (out_lod_margin) out_lod_margin[i] = fetch_margins.lod;
        if (out_margin) out_margin[i] = double(margin);
        if (out_step_margin) out_step_margin[i] = double(step_margin);
        for (int k = 0; k < 4; ++k) out_rgba[i * 4 + k] = O(o.rgba[k]);
        out_dist[i] = O(o.distance);
        if (out_steps) out_steps[i] = o.steps;
    }
}


template <class S>
PathTracer<S> make_tracer(const Atmosphere<S>& A, const RenderBlock<S>& R, const SkyhipCloudCommonBufferData* common,
                          const SkyhipCloudMaterial* material, const SkyhipPathTracingParameters* params,
                          const SkyhipPathTracingInputs* in) {
    PathTracer<S> P(A, R, *common, *material, *params);
    P.T = in->transmittance, P.tw = in->transmittance_width, P.th = in->transmittance_height;
    P.sv_l = in->sky_view_luminance, P.sv_t = in->sky_view_transmittance;
    P.sw = in->sky_view_width, P.sh = in->sky_view_height;
    P.ap_l = in->ap_luminance, P.ap_t = in->ap_transmittance;
    P.aw = in->ap_width, P.ah = in->ap_height, P.ad = in->ap_depth;
    if (in->brick_majorants && material->kind == SKYHIP_MATERIAL_VOXEL_F16)
        P.use_bricks(in->brick_majorants, in->brick_size ? in->brick_size : kBrickTexels);
    return P;
}

// Samples of the listed pixels for frames first_frame .. first_frame + frames - 1:
// out[(i * frames + f) * 4 + c].
template <class S, class O>
void path_trace(const SkyhipAtmosphereBufferData* atm, const SkyhipAtmosphereRenderBufferData* render,
                const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                const SkyhipPathTracingParameters* params, const SkyhipPathTracingInputs* in, int width, int height,
                const int32_t* pixels, long count, uint32_t first_frame, int frames, O* out) {
    Atmosphere<S> A(*atm);
    RenderBlock<S> R(*render);
    PathTracer<S> P = make_tracer<S>(A, R, common, material, params, in);
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < count; ++i) {
        int p = pixels[i];
        for (int f = 0; f < frames; ++f) {
            S o[4];
            P.sample(p % width, p / width, width, height, first_frame + uint32_t(f), o);
            for (int c = 0; c < 4; ++c) out[(i * frames + f) * 4 + c] = O(o[c]);
        }
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
// GetSunVisibility (Atmosphere.glsl:110-117) at (r, mu_s) with transmittance LUT T (RGBA32F tw x th).
void oracle_sun_visibility_f64(const SkyhipAtmosphereBufferData* atm, const float* T, int tw, int th, double r,
                               double mu_s, double* out) {
    Atmosphere<double> A(*atm);
    V3<double> v = A.GetSunVisibility(T, tw, th, r, mu_s);
    out[0] = v.x, out[1] = v.y, out[2] = v.z;
}
// The samplers the oracle emulates, at normalised coordinates: LINEAR / CLAMP_TO_EDGE of RGBA32F
// 2D and 3D textures (rgb), and textureLod of an UNORM8 2D mip chain with REPEAT.
void oracle_tex2d_linear_f64(const float* rgba, int w, int h, double u, double v, double* out) {
    V3<double> r = tex2d_linear_clamp<double>(rgba, w, h, u, v);
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}
void oracle_tex3d_linear_f64(const float* rgba, int w, int h, int d, double u, double v, double q, double* out) {
    V3<double> r = tex3d_linear_clamp<double>(rgba, w, h, d, u, v, q);
    out[0] = r.x, out[1] = r.y, out[2] = r.z;
}
void oracle_texlod2d_repeat_f64(const uint8_t* tex, int size, int levels, int channels, double u, double v,
                                double lambda, double* out) {
    texlod2d_u8_repeat<double>(tex, size, levels, channels, u, v, lambda, channels, out);
}
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


// ---------------------------------------------------------------------------
// Noise textures (NoiseGen.comp) -- level 0 over a list of texel indices, so
// large volumes can be checked on a sample.
// ---------------------------------------------------------------------------
void oracle_cloud_map_texels(const SkyhipNoiseCreateInfo* density, const SkyhipNoiseCreateInfo* height, int size,
                             const int32_t* texels, long count, uint8_t* out_rg, int f64) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < count; ++i) {
        int t = texels[i];
        if (f64)
            CloudMapTexel<double>(t % size, t / size, size, *density, *height, out_rg + 2 * i);
        else
            CloudMapTexel<float>(t % size, t / size, size, *density, *height, out_rg + 2 * i);
    }
}
void oracle_detail_texels(const SkyhipNoiseCreateInfo* perlin, const SkyhipNoiseCreateInfo* worley, int size,
                          const int32_t* texels, long count, uint8_t* out, int f64) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < count; ++i) {
        int t = texels[i];
        int x = t % size, y = (t / size) % size, z = t / (size * size);
        out[i] = f64 ? DetailTexel<double>(x, y, z, size, *perlin, *worley)
                     : DetailTexel<float>(x, y, z, size, *perlin, *worley);
    }
}
void oracle_worley_texels(const SkyhipNoiseCreateInfo* worley, int size, const int32_t* texels, long count,
                          uint8_t* out, int f64) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < count; ++i) {
        int t = texels[i];
        int x = t % size, y = (t / size) % size, z = t / (size * size);
        out[i] = f64 ? WorleyTexel<double>(x, y, z, size, *worley) : WorleyTexel<float>(x, y, z, size, *worley);
    }
}
void oracle_displacement_texels(const SkyhipNoiseCreateInfo* perlin, int size, const int32_t* texels, long count,
                                uint8_t* out_rgba, int f64) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < count; ++i) {
        int t = texels[i];
        if (f64)
            DisplacementTexel<double>(t % size, t / size, size, *perlin, out_rgba + 4 * i);
        else
            DisplacementTexel<float>(t % size, t / size, size, *perlin, out_rgba + 4 * i);
    }
}
double oracle_perlin_noise(double x, double y, double z, unsigned freq, unsigned seed) {
    return PerlinNoise<double>({x, y, z}, freq, seed);
}
double oracle_worley_noise(double x, double y, double z, unsigned freq, unsigned seed) {
    return WorleyNoise<double>({x, y, z}, freq, seed);
}
unsigned oracle_wang_hash(unsigned x) { return WangHash(x); }
unsigned oracle_pcg_hash(unsigned x) { return PCGHash(x); }

// Fills levels 1.. of a packed UNORM8 mip chain from level 0.
void oracle_generate_mipmaps(uint8_t* tex, int dims, int channels, int w, int h, int d, int levels) {
    if (dims == 2) d = 1;
    int sd[3] = {w, h, d};
    size_t off = 0;
    for (int l = 1; l < levels; ++l) {
        int dd[3] = {std::max(1, sd[0] / 2), std::max(1, sd[1] / 2), dims == 3 ? std::max(1, sd[2] / 2) : 1};
        size_t src_bytes = size_t(sd[0]) * sd[1] * sd[2] * channels;
        downsample_u8(tex + off, sd, tex + off + src_bytes, dd, dims, channels);
        off += src_bytes;
        sd[0] = dd[0], sd[1] = dd[1], sd[2] = dd[2];
    }
}

// ---------------------------------------------------------------------------
// Cloud raymarch (VolumetricCloudRender.comp) over a list of output pixels.
// out_rgba: 4 per pixel, out_dist: 1 per pixel, out_steps: executed primary
// steps per pixel (may be NULL).
// ---------------------------------------------------------------------------
void oracle_cloud_raymarch_f64(const SkyhipAtmosphereBufferData* atm, const SkyhipCloudCommonBufferData* common,
                               const SkyhipCloudBufferData* cloud, const SkyhipCloudMaterial* material,
                               const CloudImages* images, int width, const int32_t* pixels, long count,
                               double* out_rgba, double* out_dist, int32_t* out_steps) {
    cloud_raymarch<double>(atm, common, cloud, material, images, width, pixels, count, out_rgba, out_dist, out_steps);
}
void oracle_cloud_raymarch_f32(const SkyhipAtmosphereBufferData* atm, const SkyhipCloudCommonBufferData* common,
                               const SkyhipCloudBufferData* cloud, const SkyhipCloudMaterial* material,
                               const CloudImages* images, int width, const int32_t* pixels, long count,
                               float* out_rgba, float* out_dist, int32_t* out_steps) {
    cloud_raymarch<float>(atm, common, cloud, material, images, width, pixels, count, out_rgba, out_dist, out_steps);
}

// Checker mode: replays per-pixel step decisions (forced_steps[2 i], [2 i + 1] for
// pixels[i]) and reports each pixel's decision margins (see CloudRaymarchPixel).
void oracle_cloud_raymarch_forced_f64(const SkyhipAtmosphereBufferData* atm, const SkyhipCloudCommonBufferData* common,
                                      const SkyhipCloudBufferData* cloud, const SkyhipCloudMaterial* material,
                                      const CloudImages* images, int width, const int32_t* pixels, long count,
                                      const uint32_t* forced_steps, double* out_rgba, double* out_dist,
                                      int32_t* out_steps, double* out_margin, double* out_step_margin,
                                      double* out_texel_margin, double* out_lod_margin) {
    cloud_raymarch<double>(atm, common, cloud, material, images, width, pixels, count, out_rgba, out_dist, out_steps,
                           forced_steps, out_margin, out_step_margin, out_texel_margin, out_lod_margin);
}

// The FP32 instantiation of the checker mode: an independent FP32 evaluation under the same
// replayed decisions, which measures how far FP32 rounding alone moves a pixel.
void oracle_cloud_raymarch_forced_f32(const SkyhipAtmosphereBufferData* atm, const SkyhipCloudCommonBufferData* common,
                                      const SkyhipCloudBufferData* cloud, const SkyhipCloudMaterial* material,
                                      const CloudImages* images, int width, const int32_t* pixels, long count,
                                      const uint32_t* forced_steps, float* out_rgba, float* out_dist,
                                      int32_t* out_steps, double* out_margin, double* out_step_margin,
                                      double* out_texel_margin, double* out_lod_margin) {
    cloud_raymarch<float>(atm, common, cloud, material, images, width, pixels, count, out_rgba, out_dist, out_steps,
                          forced_steps, out_margin, out_step_margin, out_texel_margin, out_lod_margin);
}

// Per-step trace of one pixel (8 doubles per step, see CloudRaymarchPixel); returns the step count.
int oracle_cloud_trace_f64(const SkyhipAtmosphereBufferData* atm, const SkyhipCloudCommonBufferData* common,
                           const SkyhipCloudBufferData* cloud, const SkyhipCloudMaterial* material,
                           const CloudImages* images, int px, int py, double* trace, int trace_max) {
    Atmosphere<double> A(*atm);
    CloudCtx<double> Cc(common, cloud, material);
    return CloudRaymarchPixel<double>(A, Cc, *images, px, py, nullptr, nullptr, trace, trace_max).steps;
}

// One density evaluation (for material tests).
double oracle_cloud_sigma_t_f64(const SkyhipCloudCommonBufferData* common, const SkyhipCloudBufferData* cloud,
                                const SkyhipCloudMaterial* material, double x, double y, double z) {
    CloudCtx<double> Cc(common, cloud, material);
    V3<double> p = {x, y, z};
    return Cc.SampleSigmaT(p, Cc.CalHeight01(p));
}
double oracle_cloud_shadow_f64(const SkyhipCloudCommonBufferData* common, const SkyhipCloudBufferData* cloud,
                               const SkyhipCloudMaterial* material, double x, double y, double z) {
    CloudCtx<double> Cc(common, cloud, material);
    return Cc.SampleShadow({x, y, z});
}

// ---------------------------------------------------------------------------
// Voxel-cloud path tracer (VolumetricCloudPathTracing.comp), global majorant.
// ---------------------------------------------------------------------------
void oracle_path_trace_f64(const SkyhipAtmosphereBufferData* atm, const SkyhipAtmosphereRenderBufferData* render,
                           const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                           const SkyhipPathTracingParameters* params, const SkyhipPathTracingInputs* in, int width,
                           int height, const int32_t* pixels, long count, uint32_t first_frame, int frames,
                           double* out) {
    path_trace<double>(atm, render, common, material, params, in, width, height, pixels, count, first_frame, frames,
                       out);
}
void oracle_path_trace_f32(const SkyhipAtmosphereBufferData* atm, const SkyhipAtmosphereRenderBufferData* render,
                           const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                           const SkyhipPathTracingParameters* params, const SkyhipPathTracingInputs* in, int width,
                           int height, const int32_t* pixels, long count, uint32_t first_frame, int frames,
                           float* out) {
    path_trace<float>(atm, render, common, material, params, in, width, height, pixels, count, first_frame, frames,
                      out);
}

// Pieces of the tracer for closed-form checks.
void oracle_pt_hg_sample_f64(const SkyhipPathTracingParameters* params, uint32_t seed, const double* d, double* out_dir,
                             double* out_value_over_pdf, uint32_t* out_seed) {
    static const SkyhipAtmosphereBufferData zero_atm = {};
    static const SkyhipAtmosphereRenderBufferData zero_render = {};
    static const SkyhipCloudCommonBufferData zero_common = {};
    static const SkyhipCloudMaterial zero_material = {};
    Atmosphere<double> A(zero_atm);
    RenderBlock<double> R(zero_render);
    PathTracer<double> P(A, R, zero_common, zero_material, *params);
    double vop;
    V3<double> r = P.hg_sample(seed, {d[0], d[1], d[2]}, vop);
    out_dir[0] = r.x, out_dir[1] = r.y, out_dir[2] = r.z;
    *out_value_over_pdf = vop;
    *out_seed = seed;
}
void oracle_pt_lambert_sample_f64(uint32_t seed, const double* n, double* out_dir, uint32_t* out_seed) {
    static const SkyhipAtmosphereBufferData zero_atm = {};
    static const SkyhipAtmosphereRenderBufferData zero_render = {};
    static const SkyhipCloudCommonBufferData zero_common = {};
    static const SkyhipCloudMaterial zero_material = {};
    static const SkyhipPathTracingParameters params = {};
    Atmosphere<double> A(zero_atm);
    RenderBlock<double> R(zero_render);
    PathTracer<double> P(A, R, zero_common, zero_material, params);
    V3<double> r = P.lambert_sample(seed, {n[0], n[1], n[2]});
    out_dir[0] = r.x, out_dir[1] = r.y, out_dir[2] = r.z;
    *out_seed = seed;
}
double oracle_pt_transmittance_f64(const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                                   const SkyhipPathTracingParameters* params, uint32_t seed, const double* o,
                                   const double* d) {
    static const SkyhipAtmosphereBufferData zero_atm = {};
    static const SkyhipAtmosphereRenderBufferData zero_render = {};
    Atmosphere<double> A(zero_atm);
    RenderBlock<double> R(zero_render);
    PathTracer<double> P(A, R, *common, *material, *params);
    return P.transmittance_estimation(seed, {o[0], o[1], o[2]}, {d[0], d[1], d[2]});
}
// One tracking call from o along d (unit) through the cloud region, with the single
// majorant or (majorants != NULL, VOXEL_F16) the brick grid.  mode 0: delta tracking,
// returns the collision distance from o or -1 when the ray leaves; mode 1: ratio
// tracking, returns the transmittance estimate.
double oracle_pt_track_f64(const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                           const SkyhipPathTracingParameters* params, const float* majorants, int brick, int mode,
                           uint32_t seed, const double* o, const double* d) {
    static const SkyhipAtmosphereBufferData zero_atm = {};
    static const SkyhipAtmosphereRenderBufferData zero_render = {};
    Atmosphere<double> A(zero_atm);
    RenderBlock<double> R(zero_render);
    PathTracer<double> P(A, R, *common, *material, *params);
    if (majorants) P.use_bricks(majorants, brick);
    V3<double> oo = {o[0], o[1], o[2]}, dd = {d[0], d[1], d[2]};
    if (mode == 1) return P.transmittance_estimation(seed, oo, dd);
    double t0, t1, t;
    P.region(oo, dd, t0, t1);
    if (t0 >= t1) return -1.0;
    return P.delta_track(seed, oo, dd, t0, t1, t) ? t : -1.0;
}
double oracle_pt_sigma_t_f64(const SkyhipCloudCommonBufferData* common, const SkyhipCloudMaterial* material,
                             const SkyhipPathTracingParameters* params, double x, double y, double z) {
    static const SkyhipAtmosphereBufferData zero_atm = {};
    static const SkyhipAtmosphereRenderBufferData zero_render = {};
    Atmosphere<double> A(zero_atm);
    RenderBlock<double> R(zero_render);
    PathTracer<double> P(A, R, *common, *material, *params);
    return P.sigma({x, y, z});
}

// Sky radiance of (altitude, mu_view, mu_sun, delta_phi) queries: the sky-view texel
// arithmetic (AtmosphereRenderer.glsl:153-186) for an explicit direction, with the
// planet centre below the camera and the sun at azimuth 0.  queries: 5 doubles each
// (atmosphere index, altitude, mu_view, mu_sun, delta_phi); LUTs per atmosphere at a
// uniform stride; out: 3 doubles per query.
}  // extern "C"

template <class S>
void sky_radiance(const SkyhipAtmosphereBufferData* atms, const float* T, int tw, int th, const float* MS, int mw,
                  int mh, float steps, const double* queries, long count, S* out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < count; ++i) {
        const double* q = queries + 5 * i;
        const int a = int(q[0]);
        Atmosphere<S> A(atms[a]);
        const float* Ta = T + size_t(a) * tw * th * 4;
        const float* Ma = MS + size_t(a) * mw * mh * 4;
        S h = S(q[1]), mu_v = S(q[2]), mu_s = S(q[3]), dphi = S(q[4]);
        V3<S> c = {S(0), S(0), -A.bottom_radius};
        V3<S> start = {S(0), S(0), h};
        S sv = std::sqrt(std::fmax(S(1) - mu_v * mu_v, S(0))), ss = std::sqrt(std::fmax(S(1) - mu_s * mu_s, S(0)));
        V3<S> dir = {sv * std::cos(dphi), sv * std::sin(dphi), mu_v};
        V3<S> sun = {ss, S(0), mu_s};
        S r = A.bottom_radius + h;
        S dist = A.RayIntersectsGround(r, mu_v) ? A.DistanceToBottomAtmosphereBoundary(r, mu_v)
                                                : A.DistanceToTopAtmosphereBoundary(r, mu_v);
        V3<S> L = {S(0), S(0), S(0)}, tr;
        if (dist > S(0))
            L = A.ComputeScatteredLuminance(false, Ta, tw, th, Ma, mw, mh, 0.5f, c, start, dir, sun, dist, steps, tr,
                                            nullptr);
        out[3 * i] = L.x, out[3 * i + 1] = L.y, out[3 * i + 2] = L.z;
    }
}
extern "C" {

void oracle_sky_radiance_f64(const SkyhipAtmosphereBufferData* atms, const float* T, int tw, int th, const float* MS,
                             int mw, int mh, float steps, const double* queries, long count, double* out) {
    sky_radiance<double>(atms, T, tw, th, MS, mw, mh, steps, queries, count, out);
}
// The FP32 instantiation (the CPU baseline of the plain integrator).
void oracle_sky_radiance_f32(const SkyhipAtmosphereBufferData* atms, const float* T, int tw, int th, const float* MS,
                             int mw, int mh, float steps, const double* queries, long count, float* out) {
    sky_radiance<float>(atms, T, tw, th, MS, mw, mh, steps, queries, count, out);
}

// The FP16 density grid of the path-tracing workload, and its brick majorants.
void oracle_voxel_grid_f16(const SkyhipVoxelGridCreateInfo* ci, int dx, int dy, int dz, uint16_t* out) {
    const int dim[3] = {dx, dy, dz};
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = 0; z < dz; ++z)
        for (int y = 0; y < dy; ++y)
            for (int x = 0; x < dx; ++x)
                out[(size_t(z) * dy + y) * dx + x] = float_to_half_rne(VoxelGridDensity(x, y, z, dim, *ci));
}
void oracle_voxel_grid_density_f32(const SkyhipVoxelGridCreateInfo* ci, int dx, int dy, int dz, const int32_t* xyz,
                                   long count, float* out) {
    const int dim[3] = {dx, dy, dz};
    for (long i = 0; i < count; ++i) out[i] = VoxelGridDensity(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], dim, *ci);
}
void oracle_brick_majorants(const uint16_t* grid, int dx, int dy, int dz, int B, float* out) {
    const int dim[3] = {dx, dy, dz};
    int bx = (dx + B - 1) / B, by = (dy + B - 1) / B, bz = (dz + B - 1) / B;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < bz; ++z)
        for (int y = 0; y < by; ++y)
            for (int x = 0; x < bx; ++x) out[(size_t(z) * by + y) * bx + x] = BrickMajorant(grid, dim, x, y, z, B);
}
// Environment-luminance cube (all levels) and IBL precomputation; cubes are uint16 RGBA.
void oracle_environment_cube_f64(const SkyhipAtmosphereBufferData* atm, const SkyhipAtmosphereRenderBufferData* render,
                                 const float* T, int tw, int th, const float* sv_l, const float* sv_t, int sw, int sh,
                                 int w, uint16_t* out) {
    Atmosphere<double> A(*atm);
    RenderBlock<double> R(*render);
    EnvironmentCube<double>(A, R, T, tw, th, sv_l, sv_t, sw, sh, w, out);
}
void oracle_environment_cube_f32(const SkyhipAtmosphereBufferData* atm, const SkyhipAtmosphereRenderBufferData* render,
                                 const float* T, int tw, int th, const float* sv_l, const float* sv_t, int sw, int sh,
                                 int w, uint16_t* out) {
    Atmosphere<float> A(*atm);
    RenderBlock<float> R(*render);
    EnvironmentCube<float>(A, R, T, tw, th, sv_l, sv_t, sw, sh, w, out);
}
void oracle_cube_mips(uint16_t* cube, int w) {
    for (int l = 1; l <= Log2Int(w); ++l) CubeMipLevel(cube, w, l);
}
void oracle_env_radiance_sh_f64(const uint16_t* cube, int w, double* out) { EnvRadianceSH<double>(cube, w, out); }
void oracle_prefiltered_radiance_f64(const uint16_t* env, int ew, int pw, int count, uint16_t* out) {
    PrefilteredRadiance<double>(env, ew, pw, count, out);
}
void oracle_prefiltered_radiance_f32(const uint16_t* env, int ew, int pw, int count, uint16_t* out) {
    PrefilteredRadiance<float>(env, ew, pw, count, out);
}
void oracle_cube_sample_f64(const uint16_t* cube, int w, int levels, const double* d, double lod, double* out) {
    V3<double> c = CubeSampleLod<double>(cube, w, levels, {d[0], d[1], d[2]}, lod);
    out[0] = c.x, out[1] = c.y, out[2] = c.z;
}
void oracle_cube_texel_dir_f64(int face, int x, int y, int w, double* out) {
    V3<double> d = CubeTexelDir<double>(face, x, y, w);
    out[0] = d.x, out[1] = d.y, out[2] = d.z;
}
void oracle_cube_face_st_f64(const double* d, int* face, double* s, double* t) {
    CubeFaceST<double>({d[0], d[1], d[2]}, *face, *s, *t);
}
// Cloud shadow passes: maps are RG interleaved (depth, transmittance).
void oracle_cloud_shadow_map_f64(const SkyhipCloudCommonBufferData* c, const SkyhipCloudMaterial* m, const uint16_t* bn,
                                 int bn_size, const float* previous, int w, int h, double* out) {
    CloudShadowMap<double>(*c, *m, bn, bn_size, previous, w, h, out);
}
void oracle_cloud_shadow_map_f32(const SkyhipCloudCommonBufferData* c, const SkyhipCloudMaterial* m, const uint16_t* bn,
                                 int bn_size, const float* previous, int w, int h, float* out) {
    CloudShadowMap<float>(*c, *m, bn, bn_size, previous, w, h, out);
}
void oracle_cloud_shadow_blur(const float* in, float* out, int w, int h, int axis) { CloudShadowBlur(in, out, w, h, axis); }
void oracle_cloud_shadow_froxel_f64(const SkyhipCloudCommonBufferData* c, const float* map, int mw, int mh, int w, int h,
                                    int d, uint16_t* out) {
    CloudShadowFroxel<double>(*c, map, mw, mh, w, h, d, out);
}
void oracle_cloud_shadow_froxel_f32(const SkyhipCloudCommonBufferData* c, const float* map, int mw, int mh, int w, int h,
                                    int d, uint16_t* out) {
    CloudShadowFroxel<float>(*c, map, mw, mh, w, h, d, out);
}
void oracle_cloud_checkerboard_depth(const float* depth, int w, int h, float* out) { CheckerboardDepth(depth, w, h, out); }
void oracle_cloud_shading_index_f64(const SkyhipCloudCommonBufferData* c, const float* cb, int cw, int ch, double* out) {
    ShadingIndex<double>(*c, cb, cw, ch, out);
}
void oracle_cloud_shading_index_f32(const SkyhipCloudCommonBufferData* c, const float* cb, int cw, int ch, float* out) {
    ShadingIndex<float>(*c, cb, cw, ch, out);
}
void oracle_cloud_reconstruct_f64(const SkyhipCloudCommonBufferData* c, const float* cb, const float* index,
                                  const uint16_t* render, const float* distance, const uint16_t* previous, int w, int h,
                                  double* out) {
    Reconstruct<double>(*c, cb, index, render, distance, previous, w, h, out);
}
// FP32, stored as RGBA16F bits.
void oracle_cloud_reconstruct_f32(const SkyhipCloudCommonBufferData* c, const float* cb, const float* index,
                                  const uint16_t* render, const float* distance, const uint16_t* previous, int w, int h,
                                  uint16_t* out) {
    std::vector<float> v(size_t(w / 2) * (h / 2) * 4);
    Reconstruct<float>(*c, cb, index, render, distance, previous, w, h, v.data());
    for (size_t i = 0; i < v.size(); ++i) out[i] = float_to_half_rne(v[i]);
}
void oracle_cloud_upscale_f64(const SkyhipCloudCommonBufferData* c, const float* cb, const float* depth,
                              const uint16_t* reconstruct, const uint16_t* hdr, int w, int h, double* out) {
    Upscale<double>(*c, cb, depth, reconstruct, hdr, w, h, out);
}
void oracle_cloud_upscale_f32(const SkyhipCloudCommonBufferData* c, const float* cb, const float* depth,
                              const uint16_t* reconstruct, const uint16_t* hdr, int w, int h, uint16_t* out) {
    std::vector<float> v(size_t(w) * h * 4);
    Upscale<float>(*c, cb, depth, reconstruct, hdr, w, h, v.data());
    for (size_t i = 0; i < v.size(); ++i) out[i] = float_to_half_rne(v[i]);
}
void oracle_bind_shadow_froxel(const uint16_t* froxel, int w, int h, int d) {
    BoundFroxel() = FroxelBinding{froxel, froxel ? w : 0, froxel ? h : 0, froxel ? d : 0};
}
void oracle_bind_cloud_phase(const SkyhipCloudPhase* phase) {
    BoundPhase() = phase ? *phase : ReferencePhase();
}
uint16_t oracle_float_to_half(float x) { return float_to_half_rne(x); }
float oracle_half_to_float(uint16_t h) { return half_to_float(h); }

}  // extern "C"
