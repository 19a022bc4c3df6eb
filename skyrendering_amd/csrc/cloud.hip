// Volumetric-cloud raymarch for gfx950 (VolumetricCloudRender.comp:139-210).
//
// One ray per lane.  A wave64 covers an 8x8 pixel tile (a 256-lane workgroup
// covers 16x16), so the rays of a wave march similar step counts through
// neighbouring texels of the L2-resident cloud textures.  The march loop is
// wave-uniform: it ends when every lane of the wave has finished or dropped
// below the transmittance threshold.  Outputs: RGBA16F (luminance,
// transmittance) and R32F mean cloud distance, GL texel order.
#include "cloud_device.h"

namespace skyhip {

struct CloudArgs {
    SkyhipAtmosphereBufferData atm;
    SkyhipCloudCommonBufferData common;
    SkyhipCloudBufferData cloud;
    MaterialDev mat;
    const float4* T;
    int tw, th;
    const float4* ap_l;
    const float4* ap_t;
    int aw, ah, ad;
    const uint16_t* blue_noise;
    int bn_mask, bn_size;
    const float* depth;  // R32F at image size, or null (far plane)
    const float* index;  // RG32F shading index at output size (checkerboard mode), or null
    int image_w, image_h;
    uint16_t* out_rgba16f;
    float* out_distance;
    int width, height;
    unsigned long long* counters;  // counting variant: steps, density evals, detail fetches, cloud steps
    uint32_t* pixel_steps;  // counting variant, optional: per pixel {executed steps, scheduled steps}, seg0 | seg1 << 16
    // Light-march schedule of SampleShadow (VolumetricCloudRender.comp:98-114), evaluated once on
    // the host with the shader's FP32 counter: sample k sits at pos + D sun * light_offset[k] and
    // weighs D * delta_t[k].  light_steps > kMaxLightSteps falls back to the per-lane loop.
    int light_steps;
    float light_offset[16];
    float light_weight[16];
    // The context's bound ray-scatter-visibility froxel, or null (visibility 1).
    const uint16_t* froxel;
    int fw, fh, fd;
    SkyhipCloudPhase phase;  // the context's dual-lobe phase
};

struct MarchAcc {
    float transmittance, transmittance_sum, weighted_t_sum;
    float sun_env_r, sun_env_g;
};

constexpr int kMaxLightSteps = 16;

// The part of RayMarchStep (VolumetricCloudRender.comp:116-137) after sigma_t and the
// transmittance to the sun are known.  hg_back = HG(cos, back_g) * back_scale and
// hg_forward = HG(cos, forward_g) depend on the ray only (cos = dot(sun, view)), so they are
// evaluated once per ray.
__device__ inline void march_accumulate(const CloudArgs& a, float sigma_t, float t_sun, float t, float h01,
                                        float step_size, float hg_back, float hg_forward, MarchAcc& acc) {
    const float tr = __expf(-step_size * sigma_t);
    const float w = __expf(-a.cloud.uSunMultiscatteringSigmaScale * sigma_t);
    const float phase = hg_back * (1.0f - w) + hg_forward * w;  // mix(back, forward, w)
    float se_r = t_sun * phase;
    float se_g = a.cloud.uEnvBottomVisibility * (1.0f - h01) + h01;
    se_g = se_g - se_g * __expf(-a.cloud.uEnvMultiscatteringSigmaScale * sigma_t);
    se_r = se_r - se_r * tr;
    se_g = se_g - se_g * tr;
    acc.sun_env_r += acc.transmittance * se_r;
    acc.sun_env_g += acc.transmittance * se_g;
    acc.transmittance_sum += acc.transmittance;
    acc.weighted_t_sum += t * acc.transmittance;
    acc.transmittance *= tr;
}

// Orders LDS traffic between the lanes of one wave (no cross-wave barrier: the
// waves of a workgroup march different numbers of steps).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-lane state of one ray's march (two shell segments).
struct RayState {
    F3 view;
    float frag_dist, u, v, jitter, hg_back, hg_forward;
    float t1_0, t1_1, t2_1;
    float t_start, step_size;
    uint32_t k, n;
    int seg;
    bool done;
    uint32_t seg_steps[2], seg_sched[2];
    MarchAcc acc;

    // Segment set-up of VolumetricCloudRender.comp:159-169 / :175-181.
    __device__ void begin_segment(const CloudArgs& a, float t1, float t2, float jit, int s) {
        float dist = fminf(t2 - t1, a.cloud.uMaxRaymarchDistance);
        n = uint32_t(fmaxf(a.cloud.uMaxRaymarchSteps * (dist / a.cloud.uMaxRaymarchDistance), 1.0f));
        step_size = dist / float(n);
        // Sample k sits at t1 + (k + jitter) step_size, computed directly (no FP32 drift
        // from accumulating `t += step_size`).
        t_start = t1 + step_size * jit;
        k = 0;
        seg = s;
        seg_sched[s] = n;
    }
};

// Six waves per SIMD (79 VGPRs, no spill).  On the config-2 frame: five waves 0.493-0.495 ms,
// six 0.461-0.464 ms, seven (13 VGPRs spilled) 0.466 ms; eight spilled more and was slower.
// Every variant gives the same image.
template <int kMat, bool kCount>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void cloud_raymarch_kernel(CloudArgs a) {
    __shared__ float s_origin[4][64 * 3];
    // The light-march contributions, 64 * light_steps floats per wave, sized at launch so that
    // the workgroup's LDS does not cap the waves per CU below what its registers allow.
    extern __shared__ float s_contrib_all[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int px = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool active = px < a.width && py < a.height;
    int executed = 0;
    WorkCount wc;
    Shell S;
    {
        const SkyhipCloudCommonBufferData& cb = a.common;
        S.cam = f3(cb.uCameraPos);
        S.sun = f3(cb.uSunDirection);
        S.R = cb.uEarthRadius;
        S.hb = cb.uBottomAltitude;
        S.ht = cb.uTopAltitude;
        S.inv_thickness = 1.0f / (S.ht - S.hb);
    }
    RayState ray;
    ray.done = true;
    ray.seg_steps[0] = ray.seg_steps[1] = ray.seg_sched[0] = ray.seg_sched[1] = 0u;
    if (active) {
        const SkyhipCloudCommonBufferData& cb = a.common;

        int pcx = px, pcy = py;
        if (a.index) {  // checkerboard sample choice, IndexToOffset (VolumetricCloudCommon.glsl:42-52)
            uint32_t index = uint32_t(a.index[(py * a.width + px) * 2]);
            pcx = px * 2 + int(((index + 1u) >> 1) & 1u);
            pcy = py * 2 + int(((index + 2u) >> 1) & 1u);
        }
        const float u = (float(pcx) + 0.5f) / float(a.image_w), v = (float(pcy) + 0.5f) / float(a.image_h);
        const float depth = a.depth ? a.depth[pcy * a.image_w + pcx] : 1.0f;
        float frag_dist;
        const F3 view = unproject(cb.uInvMVP, u * 2.0f - 1.0f, v * 2.0f - 1.0f, depth * 2.0f - 1.0f, S.cam, &frag_dist);

        // RayShellIntersect (:39-71) with b = r mu and c = R'^2 - r^2 factored.
        const float r = S.cam.z + S.R;
        const float mu = view.z;
        const float b = r * mu;
        const float cb_ = (S.hb - S.cam.z) * (2.0f * S.R + S.hb + S.cam.z);
        const float ct_ = (S.ht - S.cam.z) * (2.0f * S.R + S.ht + S.cam.z);
        float nb, fb, db, nt, ft, dt;
        sphere_roots(b, cb_, nb, fb, db);
        sphere_roots(b, ct_, nt, ft, dt);
        float t1[2] = {0.0f, 0.0f}, t2[2] = {0.0f, 0.0f};
        if (S.cam.z < S.hb) {
            t1[0] = fb;
            t2[0] = ft;
        } else if (S.cam.z < S.ht) {
            if (db >= 0.0f && mu < 0.0f) {
                t2[0] = nb;
                t1[1] = fb;
                t2[1] = ft;
            } else {
                t2[0] = ft;
            }
        } else {
            if (db >= 0.0f && mu < 0.0f) {
                t1[0] = nt;
                t2[0] = nb;
                t1[1] = fb;
                t2[1] = ft;
            } else if (dt >= 0.0f && mu < 0.0f) {
                t1[0] = nt;
                t2[0] = ft;
            }
        }
        const float vis = fminf(frag_dist, a.cloud.uMaxVisibleDistance);
        for (int i = 0; i < 2; ++i) t2[i] = fminf(fmaxf(vis, t1[i]), t2[i]);

        const float cos_sun_view = dot(S.sun, view);
        const float noise = a.blue_noise[(py & a.bn_mask) * a.bn_size + (px & a.bn_mask)] * (1.0f / 65535.0f);
        float jitter = noise + cb.uFrameID * 0.61803398875f;
        jitter = jitter - floorf(jitter);
        ray.acc = MarchAcc{1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        ray.begin_segment(a, t1[0], t2[0], jitter, 0);
        ray.t1_1 = t1[1];
        ray.t2_1 = t2[1];
        ray.jitter = jitter;
        ray.hg_back = henyey_greenstein(cos_sun_view, a.phase.back_g) * a.phase.back_scale;
        ray.hg_forward = henyey_greenstein(cos_sun_view, a.phase.forward_g);
        ray.view = view;
        ray.frag_dist = frag_dist;
        ray.t1_0 = t1[0];
        ray.u = u;
        ray.v = v;
        ray.done = false;
    }

    // The march loop is wave-synchronous: every lane iterates until the whole wave is done, so
    // the light march of the lanes whose sample is inside a cloud can be spread over all 64 lanes.
    const bool coop = a.light_steps <= kMaxLightSteps;
    const F3 sun_vector = S.sun * a.cloud.uShadowDistance;  // sample_vector of SampleShadow
    while (__ballot(!ray.done) != 0) {
        float t = 0.0f, h01 = 0.0f, sigma = 0.0f;
        F3 pos = f3(0.0f);
        if (!ray.done) {
            t = fmaf(float(ray.k), ray.step_size, ray.t_start);
            pos = S.cam + ray.view * t;
            h01 = height01(S, pos);
            sigma = sample_sigma_t<kMat, kCount>(a.mat, pos, h01, S.cam, &wc);
        }
        const bool lit = !ray.done && sigma >= 1e-5f;  // RayMarchStep's early return
        const uint64_t lit_mask = __ballot(lit);
        float t_sun = 1.0f;
        if (lit_mask != 0) {
            if (coop) {
                // Stage the lit lanes' sample positions, evaluate all (ray, light step) pairs
                // across the wave, then each lit lane sums its own in the shader's order.
                const int m = __popcll(lit_mask);
                const int rank = __builtin_amdgcn_mbcnt_hi(uint32_t(lit_mask >> 32),
                                                           __builtin_amdgcn_mbcnt_lo(uint32_t(lit_mask), 0u));
                if (lit) {
                    s_origin[wave][3 * rank + 0] = pos.x;
                    s_origin[wave][3 * rank + 1] = pos.y;
                    s_origin[wave][3 * rank + 2] = pos.z;
                }
                wave_lds_sync();
                const int L = a.light_steps, total = m * L;
                const float inv_L = 1.0f / float(L);
                float* const s_contrib = s_contrib_all + wave * 64 * L;
                for (int base = 0; base < total; base += 64) {
                    const int item = base + lane;
                    if (item < total) {
                        // item / L for item < 64 * 16: exact through the float reciprocal (the
                        // quotient's distance from the next integer, >= 0.5 / L, is far above
                        // the rounding error), where the integer division takes ~15 instructions.
                        const int r = int((float(item) + 0.5f) * inv_L), ls = item - r * L;
                        const F3 o = {s_origin[wave][3 * r], s_origin[wave][3 * r + 1], s_origin[wave][3 * r + 2]};
                        const F3 sp = o + sun_vector * a.light_offset[ls];
                        s_contrib[item] =
                            sample_sigma_t<kMat, kCount>(a.mat, sp, height01(S, sp), S.cam, &wc) * a.light_weight[ls];
                    }
                }
                wave_lds_sync();
                if (lit) {
                    float od = 0.0f;
                    for (int ls = 0; ls < L; ++ls) od += s_contrib[rank * L + ls];
                    t_sun = __expf(-od);
                }
                wave_lds_sync();
            } else if (lit) {
                t_sun = sample_shadow<kMat, kCount>(a.mat, S, pos, a.cloud.uShadowSteps, a.cloud.uShadowDistance, &wc);
            }
        }
        if (lit) {
            if (kCount) wc.cloud_steps++;
            march_accumulate(a, sigma, t_sun, t, h01, ray.step_size, ray.hg_back, ray.hg_forward, ray.acc);
        }
        if (!ray.done) {
            ++executed;
            if (kCount) ++ray.seg_steps[ray.seg];
            ++ray.k;
            if (ray.acc.transmittance < 0.01f || ray.k >= ray.n) {  // kMinTransmittance or segment end
                if (ray.seg == 0 && ray.t2_1 - ray.t1_1 > 0.0f)
                    ray.begin_segment(a, ray.t1_1, ray.t2_1, ray.jitter, 1);
                else
                    ray.done = true;
            }
        }
    }

    if (active) {
        const SkyhipCloudCommonBufferData& cb = a.common;
        const AtmDev A = make_atm(a.atm);
        const MarchAcc& acc = ray.acc;
        const F3 view = ray.view;
        const float frag_dist = ray.frag_dist, u = ray.u, v = ray.v;
        const float r = S.cam.z + S.R, mu = view.z;
        const uint32_t* seg_steps = ray.seg_steps;
        const uint32_t* seg_sched = ray.seg_sched;
        float t1[1] = {ray.t1_0};
        if (kCount && a.pixel_steps) {
            a.pixel_steps[2 * (py * a.width + px)] = seg_steps[0] | (seg_steps[1] << 16);
            a.pixel_steps[2 * (py * a.width + px) + 1] = seg_sched[0] | (seg_sched[1] << 16);
        }
        const float average_t = acc.weighted_t_sum == 0.0f ? frag_dist : acc.weighted_t_sum / acc.transmittance_sum;
        a.out_distance[py * a.width + px] = average_t;
        const F3 ap = S.cam + view * average_t;
        // GetSunVisibility(transmittance_texture, pos) and SunCosTheta at the mean distance.
        float q = ap.x * ap.x + ap.y * ap.y + ap.z * (ap.z + 2.0f * S.R) + (S.R - A.rb) * (S.R + A.rb);
        const float rr = sqrtf(fmaxf(q + A.rb * A.rb, 0.0f));
        const float mu_s = dot(S.sun, F3{ap.x, ap.y, ap.z + S.R}) / rr;
        const F3 sun_vis = sun_visibility(A, a.T, a.tw, a.th, rr, q, q / (rr + A.rb), mu_s);
        const float sun_cos = clamp01(mu_s);
        F3 luminance = sun_vis * A.solar * (acc.sun_env_r * a.cloud.uSunIlluminanceScale) +
                       f3(a.cloud.uEnvColorScale) * (acc.sun_env_g * __powf(sun_cos, a.cloud.uEnvSunHeightCurveExp));
        // GetAerialPerspective (VolumetricCloudCommon.glsl:81-98)
        float t_ap = average_t;
        if (r > A.rt) {
            float disc = r * r * (mu * mu - 1.0f) + A.rt * A.rt;
            if (mu < 0.0f && disc >= 0.0f)
                t_ap -= fmaxf(-r * mu - sqrtf(fmaxf(disc, 0.0f)), 0.0f);
            else
                t_ap = 0.0f;
        }
        const float w = sqrtf(t_ap / cb.uAerialPerspectiveLutMaxDistance);
        const F3 ap_t = lut3d_trilinear(a.ap_t, a.aw, a.ah, a.ad, u * float(a.aw - 1), v * float(a.ah - 1),
                                        w * float(a.ad - 1));
        F3 ap_l = lut3d_trilinear(a.ap_l, a.aw, a.ah, a.ad, u * float(a.aw - 1), v * float(a.ah - 1),
                                  w * float(a.ad - 1));
        if (a.froxel)  // :198
            ap_l = ap_l * ray_scatter_visibility(a.froxel, a.fw, a.fh, a.fd, u, v, average_t,
                                                 cb.uInvShadowFroxelMaxDistance);
        luminance = luminance * ap_t + ap_l * (1.0f - acc.transmittance);
        luminance = luminance * (1.0f / fmaxf(1e-5f, 1.0f - acc.transmittance));
        const float fade = smoothstepf(a.cloud.uMaxVisibleDistance * 0.75f, a.cloud.uMaxVisibleDistance, t1[0]);
        const float T = acc.transmittance * (1.0f - fade) + fade;
        luminance = luminance * (1.0f - T);
        ushort4 o;
        o.x = __half_as_ushort(__float2half_rn(luminance.x));
        o.y = __half_as_ushort(__float2half_rn(luminance.y));
        o.z = __half_as_ushort(__float2half_rn(luminance.z));
        o.w = __half_as_ushort(__float2half_rn(T));
        reinterpret_cast<ushort4*>(a.out_rgba16f)[py * a.width + px] = o;
    }
    if (kCount) {
        uint32_t v[4] = {uint32_t(executed), wc.evals, wc.detail, wc.cloud_steps};
        for (int k = 0; k < 4; ++k) {
            uint32_t s = v[k];
            for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) atomicAdd(a.counters + k, (unsigned long long)s);
        }
    }
}

// Texture descriptor from the ABI material (levels packed back to back).
static int log2_exact(int x) {  // log2 of a power of two, else -1
    int l = 0;
    while ((1 << l) < x) ++l;
    return (1 << l) == x ? l : -1;
}

static void make_tex2d(Tex2D& t, const uint8_t* data, int size, int levels, int channels) {
    t.data = data;
    t.size = size;
    t.levels = levels;
    t.channels = channels;
    t.log2_size = log2_exact(size);
    uint32_t off = 0;
    for (int l = 0; l < kMaxMipLevels; ++l) {
        t.offset[l] = off;
        if (l < levels) {
            uint32_t s = uint32_t(size >> l > 0 ? size >> l : 1);
            off += s * s * uint32_t(channels);
        }
    }
}
static void make_tex3d(Tex3D& t, const uint8_t* data, const int dim[3], int levels) {
    t.data = data;
    t.dim[0] = dim[0], t.dim[1] = dim[1], t.dim[2] = dim[2];
    t.levels = levels;
    t.log2_size = (dim[0] == dim[1] && dim[1] == dim[2]) ? log2_exact(dim[0]) : -1;
    uint32_t off = 0;
    for (int l = 0; l < kMaxMipLevels; ++l) {
        t.offset[l] = off;
        if (l < levels) {
            uint32_t a = uint32_t(dim[0] >> l > 0 ? dim[0] >> l : 1), b = uint32_t(dim[1] >> l > 0 ? dim[1] >> l : 1),
                     c = uint32_t(dim[2] >> l > 0 ? dim[2] >> l : 1);
            off += a * b * c;
        }
    }
}

static bool pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

extern "C" int skyhip_mip_levels(int32_t width, int32_t height, int32_t depth);

int build_material(const SkyhipCloudMaterial* m, MaterialDev& d) {
    if (m == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "material is NULL");
    std::memset(&d, 0, sizeof(d));
    d.kind = m->kind;
    d.common = m->common;
    d.m1 = m->material1;
    d.m0 = m->material0;
    d.minimal_density = m->minimal.uDensity;
    d.voxel = m->voxel;
    auto finite = [](const void* p, size_t bytes) { return finite_block(reinterpret_cast<const float*>(p), bytes / 4); };
    switch (m->kind) {
        case SKYHIP_MATERIAL_DEFAULT0:
        case SKYHIP_MATERIAL_DEFAULT1:
        case SKYHIP_MATERIAL_DEFAULT1_WORLEY: {
            if (!finite(&m->common, sizeof(m->common)) || !finite(&m->material1, sizeof(m->material1)) ||
                !finite(&m->material0, sizeof(m->material0)))
                return fail(SKYHIP_ERR_NON_FINITE, "material block has a non-finite field");
            if (m->kind == SKYHIP_MATERIAL_DEFAULT1_WORLEY) {
                if (!finite(&m->worley, sizeof(m->worley)))
                    return fail(SKYHIP_ERR_NON_FINITE, "Worley detail block has a non-finite field");
                if (!m->worley_texture) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "Worley detail volume is NULL");
                if (!pow2(m->worley_size) || m->worley_size > 1024 || m->worley_levels < 1 ||
                    m->worley_levels > skyhip_mip_levels(m->worley_size, m->worley_size, m->worley_size))
                    return fail(SKYHIP_ERR_INVALID_SIZE, "Worley detail volume: power-of-two size <= 1024, valid levels");
                d.worley = m->worley;
                int wdim[3] = {m->worley_size, m->worley_size, m->worley_size};
                make_tex3d(d.worley_tex, m->worley_texture, wdim, m->worley_levels);
                d.worley_tex.cells = reinterpret_cast<const uint2*>(m->worley_cells);
            }
            // Cell-corner copies are read as 8-byte loads.
            if (reinterpret_cast<uintptr_t>(m->cloud_map_cells) % 8 != 0 ||
                reinterpret_cast<uintptr_t>(m->detail_cells) % 8 != 0 ||
                reinterpret_cast<uintptr_t>(m->worley_cells) % 8 != 0)
                return fail(SKYHIP_ERR_INVALID_ARGUMENT, "texture cell copies must be 8-byte aligned");
            if (!m->cloud_map || !m->detail) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "material texture is NULL");
            // RG8 texels are read as 2-byte and RGBA8 texels as 4-byte loads.
            if (reinterpret_cast<uintptr_t>(m->cloud_map) % 2 != 0 ||
                (m->kind == SKYHIP_MATERIAL_DEFAULT0 && reinterpret_cast<uintptr_t>(m->displacement) % 4 != 0))
                return fail(SKYHIP_ERR_INVALID_ARGUMENT, "cloud map must be 2-byte and displacement 4-byte aligned");
            // lambda = log2(k_lod distance) + bias = log2(distance) + (log2(k_lod) + bias), rounded once.
            auto lod_const = [&](const SkyhipSampleInfo& si) {
                return float(std::log2(double(si.k_lod)) + double(m->common.uLodBias));
            };
            d.lod_cloud_map = lod_const(m->common.uCloudMapSampleInfo);
            d.lod_detail = lod_const(m->common.uDetailSampleInfo);
            d.lod_displacement = lod_const(m->common.uDisplacementSampleInfo);
            if (m->kind == SKYHIP_MATERIAL_DEFAULT1_WORLEY) d.lod_worley = lod_const(m->worley.uWorleySampleInfo);
            if (!pow2(m->cloud_map_size) || !pow2(m->detail_size) || m->cloud_map_size > 16384 || m->detail_size > 1024)
                return fail(SKYHIP_ERR_INVALID_SIZE, "cloud map / detail sizes must be powers of two");
            if (m->cloud_map_levels < 1 || m->cloud_map_levels > skyhip_mip_levels(m->cloud_map_size, m->cloud_map_size, 1) ||
                m->detail_levels < 1 || m->detail_levels > skyhip_mip_levels(m->detail_size, m->detail_size, m->detail_size))
                return fail(SKYHIP_ERR_INVALID_SIZE, "bad mip level count");
            make_tex2d(d.cloud_map, m->cloud_map, m->cloud_map_size, m->cloud_map_levels, 2);
            d.cloud_map.cells = reinterpret_cast<const uint2*>(m->cloud_map_cells);
            int dim[3] = {m->detail_size, m->detail_size, m->detail_size};
            make_tex3d(d.detail, m->detail, dim, m->detail_levels);
            d.detail.cells = reinterpret_cast<const uint2*>(m->detail_cells);
            if (m->kind == SKYHIP_MATERIAL_DEFAULT0) {
                if (!m->displacement || !pow2(m->displacement_size) || m->displacement_size > 16384 ||
                    m->displacement_levels < 1 ||
                    m->displacement_levels > skyhip_mip_levels(m->displacement_size, m->displacement_size, 1))
                    return fail(SKYHIP_ERR_INVALID_SIZE, "bad displacement texture");
                make_tex2d(d.displacement, m->displacement, m->displacement_size, m->displacement_levels, 4);
            }
            return SKYHIP_OK;
        }
        case SKYHIP_MATERIAL_MINIMAL:
            if (!std::isfinite(m->minimal.uDensity)) return fail(SKYHIP_ERR_NON_FINITE, "non-finite density");
            return SKYHIP_OK;
        case SKYHIP_MATERIAL_VOXEL:
            if (!finite(&m->voxel, sizeof(m->voxel))) return fail(SKYHIP_ERR_NON_FINITE, "voxel block non-finite");
            if (!m->voxel_texture) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "voxel texture is NULL");
            for (int k = 0; k < 3; ++k)
                if (m->voxel_dim[k] < 1 || m->voxel_dim[k] > 4096)
                    return fail(SKYHIP_ERR_INVALID_SIZE, "voxel dims outside [1, 4096]");
            if (m->voxel_levels < 1 ||
                m->voxel_levels > skyhip_mip_levels(m->voxel_dim[0], m->voxel_dim[1], m->voxel_dim[2]))
                return fail(SKYHIP_ERR_INVALID_SIZE, "bad voxel mip level count");
            {  // the levels' offsets are 32-bit (level 0 is indexed in 64 bits)
                unsigned long long texels = 0;
                for (int l = 0; l < m->voxel_levels; ++l)
                    texels += (unsigned long long)std::max(m->voxel_dim[0] >> l, 1) * std::max(m->voxel_dim[1] >> l, 1) *
                              std::max(m->voxel_dim[2] >> l, 1);
                if (m->voxel_levels > 1 && texels > 0xffffffffull)
                    return fail(SKYHIP_ERR_INVALID_SIZE, "voxel mip chain of more than 2^32 texels");
            }
            make_tex3d(d.voxel_tex, m->voxel_texture, m->voxel_dim, m->voxel_levels);
            return SKYHIP_OK;
        case SKYHIP_MATERIAL_VOXEL_F16:
            if (!finite(&m->voxel, sizeof(m->voxel))) return fail(SKYHIP_ERR_NON_FINITE, "voxel block non-finite");
            if (!m->voxel_grid) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "voxel grid is NULL");
            for (int k = 0; k < 3; ++k) {
                if (m->voxel_grid_dim[k] < 2 || m->voxel_grid_dim[k] > 4096)
                    return fail(SKYHIP_ERR_INVALID_SIZE, "voxel grid dims outside [2, 4096]");
                d.grid.dim[k] = m->voxel_grid_dim[k];
            }
            d.grid.data = m->voxel_grid;
            return SKYHIP_OK;
        default:
            return fail(SKYHIP_ERR_INVALID_ARGUMENT, "unknown material kind %d", m->kind);
    }
}

template <bool kCount>
static int launch_raymarch(hipStream_t st, const CloudArgs& a) {
    dim3 grid(div_up(a.width, 16), div_up(a.height, 16));
    const size_t lds = a.light_steps <= kMaxLightSteps ? size_t(4) * 64 * a.light_steps * sizeof(float) : 0;
    switch (a.mat.kind) {
        case SKYHIP_MATERIAL_DEFAULT0:
            hipLaunchKernelGGL((cloud_raymarch_kernel<SKYHIP_MATERIAL_DEFAULT0, kCount>), grid, dim3(256), lds, st, a);
            break;
        case SKYHIP_MATERIAL_DEFAULT1:
            hipLaunchKernelGGL((cloud_raymarch_kernel<SKYHIP_MATERIAL_DEFAULT1, kCount>), grid, dim3(256), lds, st, a);
            break;
        case SKYHIP_MATERIAL_DEFAULT1_WORLEY:
            hipLaunchKernelGGL((cloud_raymarch_kernel<SKYHIP_MATERIAL_DEFAULT1_WORLEY, kCount>), grid, dim3(256), lds,
                               st, a);
            break;
        case SKYHIP_MATERIAL_MINIMAL:
            hipLaunchKernelGGL((cloud_raymarch_kernel<SKYHIP_MATERIAL_MINIMAL, kCount>), grid, dim3(256), lds, st, a);
            break;
        case SKYHIP_MATERIAL_VOXEL:
            hipLaunchKernelGGL((cloud_raymarch_kernel<SKYHIP_MATERIAL_VOXEL, kCount>), grid, dim3(256), lds, st, a);
            break;
        default:
            hipLaunchKernelGGL((cloud_raymarch_kernel<SKYHIP_MATERIAL_VOXEL_F16, kCount>), grid, dim3(256), lds, st,
                               a);
            break;
    }
    SKYHIP_LAUNCH_CHECK("cloud_raymarch_kernel");
    return SKYHIP_OK;
}

int validate_atmosphere(const SkyhipAtmosphereBufferData* b);
int validate_lut(const void* p, int w, int h, const char* what);
int validate_blue_noise(const uint16_t* bn, int size);

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_cloud_raymarch(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                     const SkyhipCloudCommonBufferData* common, const SkyhipCloudBufferData* cloud,
                                     const SkyhipCloudMaterial* material, const float* transmittance_rgba,
                                     int32_t transmittance_width, int32_t transmittance_height,
                                     const float* ap_luminance_rgba, const float* ap_transmittance_rgba,
                                     int32_t ap_width, int32_t ap_height, int32_t ap_depth,
                                     const uint16_t* blue_noise, int32_t blue_noise_size, const float* depth,
                                     const float* shading_index, void* out_rgba16f, float* out_distance,
                                     int32_t width, int32_t height, unsigned long long* work_counters,
                                     uint32_t* pixel_steps) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(validate_atmosphere(atmosphere));
    if (!common || !cloud) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "cloud blocks are NULL");
    if (!finite_block(reinterpret_cast<const float*>(common->uInvMVP), 16) ||
        !finite_block(common->uCameraPos, 3) || !finite_block(reinterpret_cast<const float*>(&common->uLinearDepthParam), 12) ||
        !finite_block(reinterpret_cast<const float*>(cloud), sizeof(*cloud) / 4))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud blocks have a non-finite field");
    if (!(common->uTopAltitude > common->uBottomAltitude) || !(cloud->uMaxRaymarchDistance > 0.0f) ||
        !(common->uAerialPerspectiveLutMaxDistance > 0.0f))
        return fail(SKYHIP_ERR_NON_FINITE, "degenerate cloud layer or raymarch distance");
    if (!(cloud->uShadowSteps > 0.0f && cloud->uShadowSteps <= 256.0f) ||
        !(cloud->uMaxRaymarchSteps >= 0.0f && cloud->uMaxRaymarchSteps <= 4096.0f))
        return fail(SKYHIP_ERR_INVALID_SIZE, "shadow steps must be in (0, 256], raymarch steps in [0, 4096]");
    CloudArgs a;
    std::memset(&a, 0, sizeof(a));
    SKYHIP_CHECK(build_material(material, a.mat));
    SKYHIP_CHECK(validate_lut(transmittance_rgba, transmittance_width, transmittance_height, "transmittance"));
    SKYHIP_CHECK(validate_lut(ap_luminance_rgba, ap_width, ap_height, "aerial-perspective luminance"));
    SKYHIP_CHECK(validate_lut(ap_transmittance_rgba, ap_width, ap_height, "aerial-perspective transmittance"));
    if (ap_depth < 2 || ap_depth > 1024) return fail(SKYHIP_ERR_INVALID_SIZE, "aerial-perspective depth outside [2, 1024]");
    if (blue_noise == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "blue noise is NULL");
    SKYHIP_CHECK(validate_blue_noise(blue_noise, blue_noise_size));
    if (!out_rgba16f || !out_distance) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "outputs are NULL");
    if (width < 1 || height < 1 || width > 16384 || height > 16384)
        return fail(SKYHIP_ERR_INVALID_SIZE, "image size %dx%d outside [1, 16384]", width, height);
    a.atm = *atmosphere;
    a.common = *common;
    a.cloud = *cloud;
    a.T = reinterpret_cast<const float4*>(transmittance_rgba);
    a.tw = transmittance_width, a.th = transmittance_height;
    a.ap_l = reinterpret_cast<const float4*>(ap_luminance_rgba);
    a.ap_t = reinterpret_cast<const float4*>(ap_transmittance_rgba);
    a.aw = ap_width, a.ah = ap_height, a.ad = ap_depth;
    a.blue_noise = blue_noise;
    a.bn_size = blue_noise_size, a.bn_mask = blue_noise_size - 1;
    a.depth = depth;
    a.index = shading_index;
    a.image_w = shading_index ? width * 2 : width;
    a.image_h = shading_index ? height * 2 : height;
    a.out_rgba16f = reinterpret_cast<uint16_t*>(out_rgba16f);
    a.out_distance = out_distance;
    a.width = width, a.height = height;
    a.counters = work_counters;
    a.pixel_steps = work_counters ? pixel_steps : nullptr;
    a.froxel = ctx->froxel;
    a.fw = ctx->froxel_w, a.fh = ctx->froxel_h, a.fd = ctx->froxel_d;
    a.phase = ctx->phase;
    if (a.froxel && !(common->uInvShadowFroxelMaxDistance > 0.0f && std::isfinite(common->uInvShadowFroxelMaxDistance)))
        return fail(SKYHIP_ERR_NON_FINITE, "a shadow froxel is bound but uInvShadowFroxelMaxDistance is not > 0");
    // SampleShadow's FP32 counter `for (t = 1/n; t <= 1.0; t += 1/n)`, evaluated once here with
    // the same float arithmetic the shader uses (the iteration count depends on its rounding).
    {
        const float inv = 1.0f / cloud->uShadowSteps;
        float prev = 0.0f;
        int L = 0;
        for (float t = inv; t <= 1.0f && L <= 256; t += inv) {
            const float cur = t * t;
            const float dt = cur - prev;
            if (L < kMaxLightSteps) {
                a.light_offset[L] = prev + 0.5f * dt;
                a.light_weight[L] = cloud->uShadowDistance * dt;
            }
            prev = cur;
            ++L;
        }
        a.light_steps = L;
    }
    return work_counters ? launch_raymarch<true>(ctx->stream, a) : launch_raymarch<false>(ctx->stream, a);
}

// ---------------------------------------------------------------------------
// The material plugin's SampleSigmaT at caller-given points (the contract of
// IVolumetricCloudMaterial's shader side, VolumetricCloudRender.comp:96).
// ---------------------------------------------------------------------------
namespace skyhip {
template <int kMat>
__global__ __launch_bounds__(256) void sample_sigma_t_kernel(MaterialDev mat, SkyhipCloudCommonBufferData cb,
                                                             const float* __restrict__ xyz, int n,
                                                             float* __restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Shell S;
    S.cam = f3(cb.uCameraPos);
    S.sun = f3(cb.uSunDirection);
    S.R = cb.uEarthRadius;
    S.hb = cb.uBottomAltitude;
    S.ht = cb.uTopAltitude;
    S.inv_thickness = 1.0f / (S.ht - S.hb);
    F3 p = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    out[i] = sample_sigma_t<kMat>(mat, p, height01(S, p), S.cam);
}
}  // namespace skyhip

extern "C" int skyhip_cloud_sample_sigma_t(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                           const SkyhipCloudMaterial* material, const float* positions_xyz,
                                           int32_t count, float* out_sigma_t) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!common || !positions_xyz || !out_sigma_t) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count < 0) return fail(SKYHIP_ERR_INVALID_SIZE, "negative count");
    if (count == 0) return SKYHIP_OK;
    MaterialDev m;
    SKYHIP_CHECK(build_material(material, m));
    dim3 grid(div_up(count, 256));
    switch (m.kind) {
        case SKYHIP_MATERIAL_DEFAULT0:
            hipLaunchKernelGGL(sample_sigma_t_kernel<SKYHIP_MATERIAL_DEFAULT0>, grid, dim3(256), 0, ctx->stream, m,
                               *common, positions_xyz, count, out_sigma_t);
            break;
        case SKYHIP_MATERIAL_DEFAULT1:
            hipLaunchKernelGGL(sample_sigma_t_kernel<SKYHIP_MATERIAL_DEFAULT1>, grid, dim3(256), 0, ctx->stream, m,
                               *common, positions_xyz, count, out_sigma_t);
            break;
        case SKYHIP_MATERIAL_DEFAULT1_WORLEY:
            hipLaunchKernelGGL(sample_sigma_t_kernel<SKYHIP_MATERIAL_DEFAULT1_WORLEY>, grid, dim3(256), 0, ctx->stream,
                               m, *common, positions_xyz, count, out_sigma_t);
            break;
        case SKYHIP_MATERIAL_MINIMAL:
            hipLaunchKernelGGL(sample_sigma_t_kernel<SKYHIP_MATERIAL_MINIMAL>, grid, dim3(256), 0, ctx->stream, m,
                               *common, positions_xyz, count, out_sigma_t);
            break;
        case SKYHIP_MATERIAL_VOXEL:
            hipLaunchKernelGGL(sample_sigma_t_kernel<SKYHIP_MATERIAL_VOXEL>, grid, dim3(256), 0, ctx->stream, m,
                               *common, positions_xyz, count, out_sigma_t);
            break;
        default:
            hipLaunchKernelGGL(sample_sigma_t_kernel<SKYHIP_MATERIAL_VOXEL_F16>, grid, dim3(256), 0, ctx->stream, m,
                               *common, positions_xyz, count, out_sigma_t);
            break;
    }
    SKYHIP_LAUNCH_CHECK("sample_sigma_t_kernel");
    return SKYHIP_OK;
}
