/*
 * skyhip.h -- C ABI of the MI355X (gfx950) implementation of SkyRendering's
 * data-parallel passes: the atmosphere LUT builds (transmittance,
 * multiple scattering, sky view, aerial perspective), the volumetric-cloud
 * raymarcher, the cloud-map / detail noise generators and the voxel-cloud path
 * tracer, plus two additions that have no reference counterpart (a batched LUT
 * builder and a learned sky-view surrogate).
 *
 * Conventions
 *  - Every parameter block below is byte-identical to the std140 uniform block
 *    the corresponding GLSL pass reads (sizes and offsets are static_asserted in
 *    the implementation).  Units are km and radians, as in the shaders.
 *  - Images are device pointers in GL texel order: texel (x, y, z) lives at
 *    ((z * height) + y) * width + x, channels interleaved.  RGBA32F LUTs are
 *    float[4] per texel.  RGBA16F images are IEEE binary16[4] per texel.
 *  - Mipmapped UNORM8 textures are one device allocation holding every level
 *    back to back (level 0 first); level n has max(1, size >> n) texels per axis.
 *  - Every function returns SKYHIP_OK (0) or a negative SKYHIP_ERR_* code and
 *    never throws.  skyhip_last_error() gives the message of the last failure
 *    on the calling thread.
 *  - Work is enqueued on the context's stream and is asynchronous unless a
 *    function says otherwise.  A context is thread-compatible: calls on one
 *    context must not overlap; different contexts may run concurrently.
 */
#ifndef SKYHIP_H
#define SKYHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes                                                               */
/* ------------------------------------------------------------------------- */
#define SKYHIP_OK 0
#define SKYHIP_ERR_INVALID_ARGUMENT (-1) /* null pointer, unknown enum, bad handle */
#define SKYHIP_ERR_INVALID_SIZE (-2)     /* zero / negative / too large / mismatched sizes */
#define SKYHIP_ERR_NON_FINITE (-3)       /* a parameter is NaN or infinite, or degenerate */
#define SKYHIP_ERR_DEVICE (-4)           /* HIP runtime error (no device, launch failure) */
#define SKYHIP_ERR_OUT_OF_MEMORY (-5)    /* device allocation failed */
#define SKYHIP_ERR_NOT_READY (-6)        /* required state missing (e.g. surrogate untrained) */

/* ------------------------------------------------------------------------- */
/* Parameter blocks (std140 images of the reference's uniform blocks)          */
/* ------------------------------------------------------------------------- */

/* AtmosphereBufferData, 128 bytes: shaders/SkyRendering/Atmosphere.glsl:6-32,
 * host struct src/SkyRendering/Atmosphere.cpp:21-47. */
typedef struct SkyhipAtmosphereBufferData {
    float solar_illuminance[3];
    float sun_angular_radius; /* radians */
    float rayleigh_scattering[3];
    float inv_rayleigh_exponential_distribution;
    float mie_scattering[3];
    float inv_mie_exponential_distribution;
    float mie_absorption[3];
    float ozone_center_altitude;
    float ozone_absorption[3];
    float inv_ozone_width;
    float ground_albedo[3];
    float mie_phase_g;
    float _atmosphere_padding[3];
    float multiscattering_mask;
    float bottom_radius;
    float top_radius;
    float transmittance_steps;
    float multiscattering_steps;
} SkyhipAtmosphereBufferData;

/* AtmosphereParameters (the JSON-facing struct), src/SkyRendering/Atmosphere.h:10-38.
 * Angles in degrees, lengths in km, scattering = colour * scale (per km). */
typedef struct SkyhipAtmosphereParameters {
    float solar_illuminance[3];
    float sun_angular_radius; /* degrees */
    float bottom_radius;
    float thickness;
    float ground_albedo[3];
    float rayleigh_exponential_distribution;
    float rayleigh_scattering_scale;
    float rayleigh_scattering[3];
    float mie_exponential_distribution;
    float mie_phase_g;
    float mie_scattering_scale;
    float mie_scattering[3];
    float mie_absorption_scale;
    float mie_absorption[3];
    float ozone_center_altitude;
    float ozone_width;
    float ozone_absorption_scale;
    float ozone_absorption[3];
    float transmittance_steps;
    float multiscattering_steps;
    float multiscattering_mask;
} SkyhipAtmosphereParameters;

/* AtmosphereRenderBufferData, 320 bytes: shaders/SkyRendering/AtmosphereRenderer.glsl:25-50,
 * host struct src/SkyRendering/AtmosphereRenderer.cpp:25-50.  Matrices are
 * column-major float[16] (GLSL / glm memory order). */
typedef struct SkyhipAtmosphereRenderBufferData {
    float sun_direction[3];
    float star_luminance_scale;
    float earth_center[3];
    float camera_earth_center_distance;
    float camera_position[3];
    float raymarching_steps;
    float up_direction[3];
    float sky_view_lut_steps;
    float right_direction[3];
    float aerial_perspective_lut_steps;
    float front_direction[3];
    float aerial_perspective_lut_max_distance;
    float moon_position[3];
    float moon_radius;
    float inv_view_projection[16];
    float light_view_projection[16];
    float padding00;
    float uInvShadowFroxelMaxDistance;
    float blocker_kernel_size_k;
    float pcss_size_k;
    float uCloudShadowMapMat[16];
} SkyhipAtmosphereRenderBufferData;

/* AtmosphereRenderParameters, src/SkyRendering/AtmosphereRenderer.h:12-35
 * (the fields the sky-view / aerial-perspective passes read). */
typedef struct SkyhipAtmosphereRenderParameters {
    float sun_direction_theta; /* zenith angle, degrees */
    float sun_direction_phi;   /* azimuth, degrees */
    float star_luminance_scale;
    float raymarching_steps;
    float sky_view_lut_steps;
    float aerial_perspective_lut_steps;
    float aerial_perspective_lut_max_distance;
    float camera_position[3];  /* world space, km, y up */
    float view_projection[16]; /* column-major */
} SkyhipAtmosphereRenderParameters;

/* Camera, src/Base/include/Camera.h + src/Base/src/Camera.cpp:8-37. */
typedef struct SkyhipCamera {
    float position[3];
    float yaw;   /* degrees */
    float pitch; /* degrees */
    float aspect;
    float fovy;  /* degrees */
    float z_near;
    float z_far;
} SkyhipCamera;

/* VolumetricCloudCommonBufferData, 384 bytes:
 * shaders/SkyRendering/VolumetricCloudCommon.glsl:6-28, src/SkyRendering/VolumetricCloud.cpp:11-33. */
typedef struct SkyhipCloudCommonBufferData {
    float uInvMVP[16];
    float uReprojectMat[16];
    float uLightVP[16];
    float uInvLightVP[16];
    float uShadowMapReprojectMat[16];
    float uCameraPos[3];
    uint32_t uBaseShadingIndex;
    float uLinearDepthParam[2];
    float uBottomAltitude;
    float uTopAltitude;
    float uSunDirection[3];
    float uFrameID;
    float uInvShadowFroxelMaxDistance;
    float uAerialPerspectiveLutMaxDistance;
    float uShadowFroxelMaxDistance;
    float uEarthRadius;
} SkyhipCloudCommonBufferData;

/* VolumetricCloudBufferData, 64 bytes: shaders/SkyRendering/VolumetricCloudRender.comp:17-32,
 * src/SkyRendering/VolumetricCloud.cpp:35-50. */
typedef struct SkyhipCloudBufferData {
    float uSunIlluminanceScale;
    float uMaxRaymarchDistance;
    float uMaxRaymarchSteps;
    float uMaxVisibleDistance;
    float uEnvColorScale[3];
    float uShadowSteps;
    float uSunMultiscatteringSigmaScale;
    float uEnvMultiscatteringSigmaScale;
    float uShadowDistance;
    float uEnvBottomVisibility;
    float padding_[3];
    float uEnvSunHeightCurveExp;
} SkyhipCloudBufferData;

/* VolumetricCloud's serialised fields, src/SkyRendering/VolumetricCloud.h:20-35. */
typedef struct SkyhipVolumetricCloudParameters {
    float bottom_altitude;
    float thickness;
    float max_raymarch_distance;
    float max_raymarch_steps;
    float max_visible_distance;
    float env_color[3];
    float env_color_scale;
    float sun_illuminance_scale;
    float shadow_steps;
    float shadow_distance;
    float shadow_map_max_distance;
    float shadow_froxel_max_distance;
    float sun_multiscattering_sigma_scale;
    float env_multiscattering_sigma_scale;
    float env_bottom_visibility;
    float env_sun_height_curve_exp;
} SkyhipVolumetricCloudParameters;

/* SampleInfo + VolumetricCloudMaterialCommonBufferData, 64 bytes:
 * shaders/SkyRendering/VolumetricCloudDefaultMaterialCommon.glsl:5-18,
 * src/SkyRendering/VolumetricCloudDefaultMaterial.cpp:9-22. */
typedef struct SkyhipSampleInfo {
    float bias[2];
    float frequency;
    float k_lod;
} SkyhipSampleInfo;

typedef struct SkyhipCloudMaterialCommonBufferData {
    SkyhipSampleInfo uCloudMapSampleInfo;
    SkyhipSampleInfo uDetailSampleInfo;
    SkyhipSampleInfo uDisplacementSampleInfo;
    float padding0[2];
    float uLodBias;
    float uDensity;
} SkyhipCloudMaterialCommonBufferData;

/* VolumetricCloudDefaultMaterial1 block, 32 bytes:
 * shaders/SkyRendering/VolumetricCloudDefaultMaterial1.glsl:3-12,
 * src/SkyRendering/VolumetricCloudDefaultMaterial.cpp:228-237. */
typedef struct SkyhipCloudMaterial1BufferData {
    float uBaseDensityThreshold;
    float uBaseHeightHardness;
    float uBaseEdgeHardness;
    float uDetailBase;
    float uDetailScale;
    float uHeightCut; /* = 1 - height_cut_ (VolumetricCloudDefaultMaterial.cpp:253) */
    float uEdgeCur;
    float padding1;
} SkyhipCloudMaterial1BufferData;

/* VolumetricCloudDefaultMaterial0 block, 16 bytes:
 * shaders/SkyRendering/VolumetricCloudDefaultMaterial0.glsl:3-7, VolumetricCloudDefaultMaterial.cpp:183-187. */
typedef struct SkyhipCloudMaterial0BufferData {
    float uDetailParam[2];
    float uDisplacementScale;
    float padding1;
} SkyhipCloudMaterial0BufferData;

/* Minimal material block, 16 bytes: shaders/SkyRendering/VolumetricCloudMaterialMinimal.glsl:1-4. */
typedef struct SkyhipCloudMinimalBufferData {
    float padding[3];
    float uDensity;
} SkyhipCloudMinimalBufferData;

/* Voxel material block, 32 bytes: shaders/SkyRendering/VolumetricCloudMaterialVoxel.glsl:3-10,
 * src/SkyRendering/VolumetricCloudVoxelMaterial.cpp:17-24. */
typedef struct SkyhipCloudVoxelBufferData {
    float uSampleFrequency[2];
    float uLodBias;
    float uDensity;
    float uSampleBias[2];
    float uSampleLodK;
    float voxel_material_padding;
} SkyhipCloudVoxelBufferData;

/* The material plugin (IVolumetricCloudMaterial, src/SkyRendering/IVolumetricCloudMaterial.h:9-30)
 * becomes a tagged parameter set: the kernels are specialised per kind, so
 * there is no virtual dispatch on the device. */
typedef enum SkyhipCloudMaterialKind {
    SKYHIP_MATERIAL_DEFAULT0 = 0, /* VolumetricCloudDefaultMaterial0.glsl */
    SKYHIP_MATERIAL_DEFAULT1 = 1, /* VolumetricCloudDefaultMaterial1.glsl */
    SKYHIP_MATERIAL_MINIMAL = 2,  /* VolumetricCloudMaterialMinimal.glsl */
    SKYHIP_MATERIAL_VOXEL = 3,    /* VolumetricCloudMaterialVoxel.glsl (R8 3D texture with mips) */
    /* The voxel material's SampleSigmaT on an FP16 density grid (level 0 only,
     * LINEAR, CLAMP_TO_BORDER(0)); the voxel block's uLodBias and uSampleLodK
     * are ignored.  No reference counterpart: it stands in for the OpenVDB-loaded
     * volume (VolumetricCloudVoxelMaterial.cpp:39-75). */
    SKYHIP_MATERIAL_VOXEL_F16 = 4,
    /* Material1 with a second, high-frequency detail volume: a WorleyFBM R8 volume
     * (skyhip_noise_worley, e.g. 32^3) that erodes Material1's density once more.  No
     * reference counterpart (Material1 samples one 128^3 detail volume,
     * VolumetricCloudDefaultMaterial1.glsl:14-28).  SampleSigmaT:
     *   d1    = clamp(base - detail, 0, 1)            Material1's density before its
     *                                                 uDensity * height01 factor
     *   mask  = max(clamp(height01 - uHeightCut, 0, 1), clamp(uEdgeCur - coverage, 0, 1))
     *                                                 Material1's erosion mask
     *   w     = (worley(GetUVWLod(pos, uWorleySampleInfo)) + uWorleyBase) * uWorleyScale * mask
     *   sigma = clamp(d1 - w, 0, 1) * uDensity * height01
     * Material1 returns 0 where its base density is 0 before any detail fetch; so does this. */
    SKYHIP_MATERIAL_DEFAULT1_WORLEY = 5
} SkyhipCloudMaterialKind;

/* The Worley detail term of SKYHIP_MATERIAL_DEFAULT1_WORLEY, 32 bytes (std140-compatible:
 * the SampleInfo of VolumetricCloudDefaultMaterialCommon.glsl:5-10, then two floats). */
typedef struct SkyhipCloudWorleyDetailBufferData {
    SkyhipSampleInfo uWorleySampleInfo;
    float uWorleyBase;
    float uWorleyScale;
    float padding[2];
} SkyhipCloudWorleyDetailBufferData;

typedef struct SkyhipCloudMaterial {
    int32_t kind; /* SkyhipCloudMaterialKind */
    SkyhipCloudMaterialCommonBufferData common;  /* DEFAULT0 / DEFAULT1 */
    SkyhipCloudMaterial1BufferData material1;    /* DEFAULT1 */
    SkyhipCloudMaterial0BufferData material0;    /* DEFAULT0 */
    SkyhipCloudMinimalBufferData minimal;        /* MINIMAL */
    SkyhipCloudVoxelBufferData voxel;            /* VOXEL */
    /* Device textures, all levels packed (see conventions).  Sampler state is
     * fixed to the reference's: REPEAT + LINEAR mag + NEAREST_MIPMAP_NEAREST min
     * for the default materials (VolumetricCloudDefaultMaterial.cpp:112-116),
     * CLAMP_TO_BORDER(0) for the voxel material (VolumetricCloudVoxelMaterial.cpp:30-37). */
    const uint8_t* cloud_map;    /* RG8 2D, cloud_map_size^2 (a power of two <= 16384) */
    int32_t cloud_map_size;
    int32_t cloud_map_levels;
    const uint8_t* detail;       /* R8 3D, detail_size^3 (a power of two <= 1024) */
    int32_t detail_size;
    int32_t detail_levels;
    const uint8_t* displacement; /* RGBA8 2D, displacement_size^2 (a power of two <= 16384; DEFAULT0 only) */
    int32_t displacement_size;
    int32_t displacement_levels;
    const uint8_t* voxel_texture; /* R8 3D, voxel_dim[0] x voxel_dim[1] x voxel_dim[2], each in [1, 4096],
                                   * with more than one level at most 2^32 texels over the levels (VOXEL only) */
    int32_t voxel_dim[3];
    int32_t voxel_levels;
    const uint16_t* voxel_grid;   /* FP16 3D, x fastest, voxel_grid_dim[0..2] (VOXEL_F16 only) */
    int32_t voxel_grid_dim[3];
    int32_t voxel_grid_padding;
    /* DEFAULT1_WORLEY only: the Worley detail block and its R8 3D volume (worley_size^3,
     * all levels packed, REPEAT + LINEAR mag + NEAREST_MIPMAP_NEAREST min like the detail
     * volume).  Appended last, so the fields above keep their offsets. */
    SkyhipCloudWorleyDetailBufferData worley;
    const uint8_t* worley_texture;
    int32_t worley_size;
    int32_t worley_levels;
    /* Optional cell-corner copies of level 0 (skyhip_texture_cells_2d_rg8 / _3d_r8), NULL to
     * read texels one by one: a LINEAR (magnification) fetch then takes one 8-byte load.  Same
     * texels and arithmetic: results are identical with and without them.  8-byte aligned. */
    const uint8_t* cloud_map_cells; /* cloud_map_size^2 cells of 8 B */
    const uint8_t* detail_cells;    /* detail_size^3 cells of 8 B */
    const uint8_t* worley_cells;    /* worley_size^3 cells of 8 B (DEFAULT1_WORLEY) */
} SkyhipCloudMaterial;

/* NoiseCreateInfo, src/SkyRendering/VolumetricCloudDefaultMaterial.h:57-62 and
 * shaders/SkyRendering/NoiseGen.comp:5-10. */
typedef struct SkyhipNoiseCreateInfo {
    uint32_t seed;
    uint32_t base_frequency;
    float remap_min;
    float remap_max;
} SkyhipNoiseCreateInfo;

/* ------------------------------------------------------------------------- */
/* Context, memory and timing helpers                                         */
/* ------------------------------------------------------------------------- */
typedef struct SkyhipContext SkyhipContext;

/* Creates a context on HIP device `device` with its own non-blocking stream. */
int skyhip_create(int device, SkyhipContext** out_ctx);
int skyhip_destroy(SkyhipContext* ctx);
/* Message of the last failed call on this thread ("" if none). */
const char* skyhip_last_error(void);
/* The context's hipStream_t, as an opaque pointer. */
void* skyhip_stream(SkyhipContext* ctx);
int skyhip_synchronize(SkyhipContext* ctx);

int skyhip_malloc(SkyhipContext* ctx, size_t bytes, void** out_device_ptr);
int skyhip_free(SkyhipContext* ctx, void* device_ptr);
int skyhip_memcpy_h2d(SkyhipContext* ctx, void* dst_device, const void* src_host, size_t bytes);
int skyhip_memcpy_d2h(SkyhipContext* ctx, void* dst_host, const void* src_device, size_t bytes);
int skyhip_memcpy_d2d(SkyhipContext* ctx, void* dst_device, const void* src_device, size_t bytes);
int skyhip_memset(SkyhipContext* ctx, void* dst_device, int value, size_t bytes);

/* Stream-ordered timing: records event `slot` (0..63) on the context's stream. */
int skyhip_event_record(SkyhipContext* ctx, int slot);
/* Milliseconds between two recorded events (synchronises on `end_slot`). */
int skyhip_event_elapsed_ms(SkyhipContext* ctx, int start_slot, int end_slot, float* out_ms);

/* Device properties used to size launches (CU count, wave size, clock). */
int skyhip_device_info(SkyhipContext* ctx, int32_t* cu_count, int32_t* wave_size, int32_t* clock_khz,
                       char* name, int32_t name_len);

/* ------------------------------------------------------------------------- */
/* Host-side parameter conversion (no device work)                            */
/* ------------------------------------------------------------------------- */

/* AssignBufferData, src/SkyRendering/Atmosphere.cpp:49-70. */
int skyhip_atmosphere_assign_buffer_data(const SkyhipAtmosphereParameters* parameters,
                                         SkyhipAtmosphereBufferData* out);

/* AssignBufferData, src/SkyRendering/AtmosphereRenderer.cpp:52-83, with
 * Earth::center() = (0, -bottom_radius, 0) (src/SkyRendering/Earth.h:37).
 * Moon, shadow-map and cloud-shadow fields are zeroed (those passes are not
 * part of this library). */
int skyhip_atmosphere_render_assign_buffer_data(const SkyhipAtmosphereRenderParameters* parameters,
                                                float bottom_radius,
                                                SkyhipAtmosphereRenderBufferData* out);

/* Camera::ViewMatrix() * Camera::ProjectionMatrix() as one column-major
 * view-projection matrix (src/Base/src/Camera.cpp:13-37, glm::lookAt and
 * glm::perspective with GL clip conventions). Also returns the camera basis. */
int skyhip_camera_view_projection(const SkyhipCamera* camera, float out_view_projection[16],
                                  float out_front[3], float out_right[3], float out_up[3]);

/* VolumetricCloud::Update, src/SkyRendering/VolumetricCloud.cpp:188-271: the
 * local z-up frame at the ground point below the camera, uInvMVP, the local
 * sun direction and camera position, and both std140 blocks.  frame_id is the
 * reference's frame_id_ (0..255).  Shadow-map matrices are left identity. */
int skyhip_cloud_assign_buffer_data(const SkyhipVolumetricCloudParameters* cloud,
                                    const SkyhipCamera* camera, float bottom_radius,
                                    const float sun_direction_world[3],
                                    float aerial_perspective_lut_max_distance, int32_t frame_id,
                                    SkyhipCloudCommonBufferData* out_common,
                                    SkyhipCloudBufferData* out_cloud,
                                    float out_model[16]);

/* The shadow-map matrices of VolumetricCloud::Update (VolumetricCloud.cpp:152-172, 219-237),
 * written into `common` (uLightVP, uInvLightVP, uShadowMapReprojectMat): the light view looks
 * from the local camera along -sun; the orthographic projection bounds the camera and its
 * frustum's far-plane corners at shadow_map_max_distance in light space, padded by half a
 * shadow-map texel.  uShadowMapReprojectMat = previous light VP * (previous model^-1 * model)
 * * translate(additional_delta) * light VP^-1 (identity when previous_light_vp is NULL;
 * previous_model goes with it; additional_delta may be NULL).  out_light_vp and
 * out_light_vp_inv_model (light VP * model^-1, the atmosphere renderer's uCloudShadowMapMat,
 * :277) may be NULL. */
int skyhip_cloud_shadow_matrices(const SkyhipVolumetricCloudParameters* cloud, const SkyhipCamera* camera,
                                 float bottom_radius, const float sun_direction_world[3],
                                 int32_t shadow_map_width, int32_t shadow_map_height,
                                 const float previous_light_vp[16], const float previous_model[16],
                                 const float additional_delta[2], SkyhipCloudCommonBufferData* common,
                                 float out_light_vp[16], float out_light_vp_inv_model[16]);

/* gen_sample_info + CalKLod, src/SkyRendering/VolumetricCloudDefaultMaterial.cpp:24-28,88-92. */
int skyhip_cloud_sample_info(float repeat_size, int32_t tex_max_width, float fovy_degrees,
                             float viewport_w, float viewport_h, double offset_x, double offset_y,
                             SkyhipSampleInfo* out);

/* ------------------------------------------------------------------------- */
/* Atmosphere LUTs                                                            */
/* ------------------------------------------------------------------------- */

/* Transmittance LUT (shaders/SkyRendering/Atmosphere.glsl:309-337), replaces the
 * TRANSMITTANCE_COMPUTE_PROGRAM dispatch in Atmosphere::UpdateLuts
 * (src/SkyRendering/Atmosphere.cpp:112-113).  out: RGBA32F width x height, a = 1. */
int skyhip_transmittance_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                             float* transmittance_rgba, int32_t width, int32_t height);

/* Multiple-scattering LUT (Atmosphere.glsl:342-438), replaces the
 * MULTISCATTERING_COMPUTE_PROGRAM dispatch (Atmosphere.cpp:116-121).  Reads the
 * transmittance LUT with bilinear CLAMP_TO_EDGE filtering.  out: RGBA32F, a = 1. */
int skyhip_multiscattering_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                               const float* transmittance_rgba, int32_t transmittance_width,
                               int32_t transmittance_height, float* multiscattering_rgba,
                               int32_t width, int32_t height);

/* Atmosphere::UpdateLuts (src/SkyRendering/Atmosphere.cpp:101-124): converts
 * the parameters and builds both LUTs, stream-ordered (the reference's missing
 * barrier between the two dispatches is a stream dependency here). */
int skyhip_atmosphere_update_luts(SkyhipContext* ctx, const SkyhipAtmosphereParameters* parameters,
                                  float* transmittance_rgba, int32_t transmittance_width,
                                  int32_t transmittance_height, float* multiscattering_rgba,
                                  int32_t multiscattering_width, int32_t multiscattering_height);

/* Sky-view LUT (shaders/SkyRendering/AtmosphereRenderer.glsl:147-186), replaces
 * sky_view_program_.Dispatch (src/SkyRendering/AtmosphereRenderer.cpp:221-226).
 * The compile-time SKY_VIEW_LUT_SIZE becomes width x height.  blue_noise may be
 * NULL (DITHER_SAMPLE_POINT_ENABLE 0, the reference default); otherwise it is an
 * R16 UNORM tile of blue_noise_size^2 (power of two) read at (x, y) & (size-1).
 * Outputs: RGBA32F (L, 0) and (T, 0). */
int skyhip_sky_view_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                        const SkyhipAtmosphereRenderBufferData* render,
                        const float* transmittance_rgba, int32_t transmittance_width,
                        int32_t transmittance_height, const float* multiscattering_rgba,
                        int32_t multiscattering_width, int32_t multiscattering_height,
                        const uint16_t* blue_noise, int32_t blue_noise_size,
                        float* luminance_rgba, float* sky_transmittance_rgba, int32_t width,
                        int32_t height);

/* Aerial-perspective LUT (AtmosphereRenderer.glsl:191-243), replaces
 * aerial_perspective_program_.Dispatch (AtmosphereRenderer.cpp:228-233).
 * Outputs: RGBA32F 3D width x height x depth, (L, 0) and (T, 0). */
int skyhip_aerial_perspective_lut(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                  const SkyhipAtmosphereRenderBufferData* render,
                                  const float* transmittance_rgba, int32_t transmittance_width,
                                  int32_t transmittance_height, const float* multiscattering_rgba,
                                  int32_t multiscattering_width, int32_t multiscattering_height,
                                  const uint16_t* blue_noise, int32_t blue_noise_size,
                                  float* luminance_rgba, float* ap_transmittance_rgba,
                                  int32_t width, int32_t height, int32_t depth);



/* One frame's atmosphere LUT set, in the reference's order: Atmosphere::UpdateLuts
 * (transmittance then multiple scattering, Atmosphere.cpp:101-124) followed by the LUT
 * passes of AtmosphereRenderer::Render (sky view and aerial perspective,
 * AtmosphereRenderer.cpp:221-233), the last two as one launch.  Built for latency: each
 * texel's march is split over several lanes and the parts combined in order, so the FP32
 * results agree with the per-pass entry points to rounding, not bit for bit.  Three kernel
 * launches on the context's stream.  Dithering off (the reference default). */
int skyhip_atmosphere_lut_set(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                              const SkyhipAtmosphereRenderBufferData* render, float* transmittance_rgba,
                              int32_t transmittance_width, int32_t transmittance_height,
                              float* multiscattering_rgba, int32_t multiscattering_width,
                              int32_t multiscattering_height, float* sky_view_luminance_rgba,
                              float* sky_view_transmittance_rgba, int32_t sky_view_width, int32_t sky_view_height,
                              float* ap_luminance_rgba, float* ap_transmittance_rgba, int32_t ap_width,
                              int32_t ap_height, int32_t ap_depth);

/* ------------------------------------------------------------------------- */
/* Environment-luminance cube and image-based lighting                       */
/* ------------------------------------------------------------------------- */

/* Cube maps are RGBA16F (4 x uint16 binary16 per texel) with a full mip chain stored level
 * by level: level l holds faces +X -X +Y -Y +Z -Z, each w_l x w_l texels row-major with
 * w_l = width >> l, texel (x, y) at face uv ((x + 0.5) / w_l, 1 - (y + 0.5) / w_l) as the
 * reference's passes write it.  Returns the texel count of the first `levels` levels
 * (levels = log2(width) + 1 for the whole chain). */
long long skyhip_cube_texels(int32_t width, int32_t levels);

/* The environment-luminance pass of AtmosphereRenderer::Render and its mip chain
 * (AtmosphereRenderer.cpp:234-242, AtmosphereRenderer.glsl:248-275): each texel is the
 * sky-view luminance along its direction, plus the lit ground times the sky-view
 * transmittance where the direction hits the ground; glGenerateTextureMipmap is a 2x2 box.
 * width: a power of two (the reference uses 128); out_cube: skyhip_cube_texels(width,
 * log2(width) + 1) texels. */
int skyhip_environment_luminance(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                                 const SkyhipAtmosphereRenderBufferData* render, const float* transmittance_rgba,
                                 int32_t transmittance_width, int32_t transmittance_height,
                                 const float* sky_view_luminance_rgba, const float* sky_view_transmittance_rgba,
                                 int32_t sky_view_width, int32_t sky_view_height, void* out_cube_rgba16f,
                                 int32_t width);

/* IBL::Precompute (src/Base/src/IBL.cpp:25-45): the 9 spherical-harmonic radiance
 * coefficients of the environment cube (EnvRadianceSH.comp; out_sh9_rgba: device vec4[9],
 * a = 0) and the GGX-prefiltered radiance cube (PrefilterRadiance.comp) with roughness_count
 * levels, level l at roughness l / (roughness_count - 1) and width prefiltered_width >> l
 * (the reference: 128 and 5); widths are powers of two <= 4096 and roughness_count lies in
 * [1, log2(prefiltered_width) + 1], else SKYHIP_ERR_INVALID_SIZE.  environment_cube: the full mip chain of
 * skyhip_environment_luminance; out_prefiltered: skyhip_cube_texels(prefiltered_width,
 * roughness_count) texels.  Lookups filter bilinearly within a face and linearly between
 * levels (the anisotropy sampler at an explicit level of detail). */
int skyhip_ibl_precompute(SkyhipContext* ctx, const void* environment_cube_rgba16f, int32_t environment_width,
                          float* out_sh9_rgba, void* out_prefiltered_rgba16f, int32_t prefiltered_width,
                          int32_t roughness_count);

/* ------------------------------------------------------------------------- */
/* Cloud shadow map and shadow froxel                                        */
/* ------------------------------------------------------------------------- */

/* VolumetricCloud::RenderShadow's shadow-map passes (VolumetricCloud.cpp:282-316):
 * VolumetricCloudShadowMap.comp into out_raw (RG32F width x height: mid-shell depth along
 * -sun, transmittance), blended 1:4 into previous_raw reprojected by uShadowMapReprojectMat
 * when previous_raw is not NULL (the reference keeps the previous frame's raw map), then the
 * 9-tap Gaussian of VolumetricCloudShadowMapBlur.comp along x into scratch and along y into
 * out_blurred.  The march samples the material's SampleSigmaT; blue_noise: the 64 x 64 (or
 * larger) UNORM16 tile, read at (x & 63, y & 63).  The reference's map is 512 x 512. */
int skyhip_cloud_shadow_map(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                            const SkyhipCloudMaterial* material, const uint16_t* blue_noise, int32_t blue_noise_size,
                            const float* previous_raw_rg32f, float* out_raw_rg32f, float* scratch_rg32f,
                            float* out_blurred_rg32f, int32_t width, int32_t height);

/* VolumetricCloudShadowFroxel.comp (VolumetricCloud.cpp:317-324): the ray-scatter visibility
 * froxel, R16 UNORM [depth][height][width] (the reference: viewport / 12 x 128 slices over
 * uShadowFroxelMaxDistance), the running mean along each view ray of
 * SampleCloudShadowTransmittance of the blurred map (LINEAR, CLAMP_TO_BORDER (1e10, 1)). */
int skyhip_cloud_shadow_froxel(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                               const float* shadow_map_rg32f, int32_t map_width, int32_t map_height,
                               uint16_t* out_froxel_r16, int32_t width, int32_t height, int32_t depth);

/* Binds a ray-scatter-visibility froxel (R16 UNORM [depth][height][width], device memory, e.g.
 * skyhip_cloud_shadow_froxel's output) to the context, as VolumetricCloud::Render and
 * PathTracing::Render bind GetShadowFroxel() (VolumetricCloud.cpp:369, 543): later
 * skyhip_cloud_raymarch and skyhip_cloud_path_trace calls on this context multiply their
 * aerial-perspective luminance by SampleRayScatterVisibility (VolumetricCloudShadowInterface.glsl:10-13),
 * LINEAR / CLAMP_TO_EDGE at (uv, distance * uInvShadowFroxelMaxDistance).  NULL unbinds
 * (visibility 1, the default).  The memory must stay valid while bound. */
int skyhip_bind_shadow_froxel(SkyhipContext* ctx, const uint16_t* froxel_r16, int32_t width, int32_t height,
                              int32_t depth);

/* The raymarch's dual-lobe phase, RayMarchStep (VolumetricCloudRender.comp:123-127):
 *   phase = mix(HG(cos, back_g) * back_scale, HG(cos, forward_g), exp(-uSunMultiscatteringSigmaScale * sigma_t))
 * The reference's lobes are the constants back_g = -0.15, back_scale = 2.16, forward_g = 0.85. */
typedef struct SkyhipCloudPhase {
    float back_g;
    float back_scale;
    float forward_g;
    float padding;
} SkyhipCloudPhase;

/* Sets the lobes of later skyhip_cloud_raymarch calls on this context (the reference's are
 * compile-time constants; this makes them a parameter).  NULL restores the reference's lobes
 * (the default).  |g| must be below 1 and back_scale finite. */
int skyhip_bind_cloud_phase(SkyhipContext* ctx, const SkyhipCloudPhase* phase);

/* ------------------------------------------------------------------------- */
/* Cloud checkerboard, shading index, temporal reconstruction and upscale     */
/* ------------------------------------------------------------------------- */

/* uReprojectMat of VolumetricCloud::Update (VolumetricCloud.cpp:204-214, 234): previous MVP *
 * translate(delta_local.xy), delta_local = inverse(model)'s rotation applied to camera
 * position - previous_camera_position (world), plus additional_delta (the material's offset
 * of this frame; may be NULL).  previous_mvp and previous_camera_position are both NULL on the
 * first frame: uReprojectMat is then this frame's MVP.  out_mvp (may be NULL) receives this
 * frame's camera view-projection * model, to pass as previous_mvp next frame.  Only
 * common->uReprojectMat is written. */
int skyhip_cloud_reproject_matrix(const SkyhipCamera* camera, float bottom_radius, const float previous_mvp[16],
                                  const float previous_camera_position[3], const float additional_delta[2],
                                  SkyhipCloudCommonBufferData* common, float out_mvp[16]);

/* CheckerboardGen.comp (VolumetricCloud.cpp:330-341): out (R32F, width/2 x height/2) = the
 * max of each 2x2 block of depth (R32F width x height) where the block's x and y have the
 * same parity, its min otherwise.  width and height: multiples of 4 in [4, 16384]. */
int skyhip_cloud_checkerboard_depth(SkyhipContext* ctx, const float* depth, int32_t width, int32_t height,
                                    float* out_checkerboard);

/* VolumetricCloudIndexGen.comp (VolumetricCloud.cpp:344-354): per 2x2 tile of the checkerboard
 * depth, the texel the quarter-resolution raymarch shades, (index, linear depth) as RG32F
 * (width/4 x height/4); index i is the texel ((i+1)>>1 & 1, (i+2)>>1 & 1) of the tile.  The
 * nearest texel wins where no neighbour tile's scheduled texel is within 25% of its linear
 * depth, then the farthest, else the rotating pattern (uBaseShadingIndex + x + y) & 3.
 * Neighbour fetches outside the checkerboard read depth 0.  width, height: the viewport. */
int skyhip_cloud_shading_index(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                               const float* checkerboard, int32_t width, int32_t height,
                               float* out_index_linear_depth);

/* VolumetricCloudReconstruct.comp (VolumetricCloud.cpp:386-404): the half-resolution temporal
 * reconstruction (RGBA16F width/2 x height/2) from the quarter-resolution render (RGBA16F)
 * and cloud distance (R32F, skyhip_cloud_raymarch with the shading index), the previous
 * frame's reconstruction reprojected by uReprojectMat (LINEAR, CLAMP_TO_EDGE) and clamped to
 * the colour box of the depth-similar rendered neighbours, blended 4:1 with the new sample
 * where this pixel was shaded, replaced by it where the reprojection leaves the screen
 * (Reinhard space).  previous_rgba16f NULL (the first frame) counts as off-screen; it must
 * not be the output. */
int skyhip_cloud_reconstruct(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                             const float* checkerboard, const float* index_linear_depth, const void* render_rgba16f,
                             const float* cloud_distance, const void* previous_rgba16f, void* out_rgba16f,
                             int32_t width, int32_t height);

/* VolumetricCloudUpscale.comp (VolumetricCloud.cpp:407-417): the depth-aware 2x upscale of
 * the reconstruction (mean of the 4 diagonal half-resolution neighbours, or the
 * depth-nearest one across a depth edge with contrasting transmittance) composited into
 * hdr (RGBA16F width x height, in place): rgb = rgb * (T <= 0.01 ? 0 : T) + luminance. */
int skyhip_cloud_upscale(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common, const float* checkerboard,
                         const float* depth, const void* reconstruct_rgba16f, void* hdr_rgba16f, int32_t width,
                         int32_t height);

/* skyhip_cloud_checkerboard_depth then skyhip_cloud_shading_index (VolumetricCloud.cpp:330-354)
 * in one launch, with the same outputs bit for bit.  depth: 16-byte aligned; the outputs
 * 8-byte aligned. */
int skyhip_cloud_checkerboard_index(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                    const float* depth, int32_t width, int32_t height, float* out_checkerboard,
                                    float* out_index_linear_depth);

/* skyhip_cloud_reconstruct then skyhip_cloud_upscale (VolumetricCloud.cpp:386-417) in one
 * launch, with the same outputs bit for bit: out_rgba16f receives the reconstruction (the next
 * frame's previous_rgba16f) and hdr_rgba16f is composited in place. */
int skyhip_cloud_reconstruct_upscale(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                     const float* checkerboard, const float* index_linear_depth,
                                     const void* render_rgba16f, const float* cloud_distance,
                                     const void* previous_rgba16f, void* out_rgba16f, const float* depth,
                                     void* hdr_rgba16f, int32_t width, int32_t height);

/* ------------------------------------------------------------------------- */
/* Cloud textures: noise generators and mip chains                           */
/* ------------------------------------------------------------------------- */

/* GetMipmapLevels, src/Base/include/ImageLoader.h:8-14: floor(log2(max dim)) + 1. */
int skyhip_mip_levels(int32_t width, int32_t height, int32_t depth);
/* Bytes of a packed mip chain (levels back to back, level n = max(1, dim >> n)). */
size_t skyhip_mip_chain_bytes(int32_t width, int32_t height, int32_t depth, int32_t channels, int32_t levels);

/* NoiseGen.comp CLOUD_MAP_GEN (shaders/SkyRendering/NoiseGen.comp:52-67), the
 * cloud_map_ DynamicTexture of VolumetricCloudDefaultMaterial.cpp:31-43: level 0
 * of a size^2 RG8 texture (R = Perlin-FBM coverage, G = Worley-FBM type). */
int skyhip_noise_cloud_map(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* density,
                           const SkyhipNoiseCreateInfo* height, uint8_t* out_rg8, int32_t size);
/* NoiseGen.comp DETAIL_MAP_GEN (:69-86), detail_ (VolumetricCloudDefaultMaterial.cpp:45-57):
 * level 0 of a size^3 R8 Perlin-Worley volume. */
int skyhip_noise_detail(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* perlin,
                        const SkyhipNoiseCreateInfo* worley, uint8_t* out_r8, int32_t size);
/* NoiseGen.comp DISPLACEMENT_GEN (:88-101), displacement_ (:59-71): level 0 of a
 * size^2 RGBA8 texture. */
int skyhip_noise_displacement(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* perlin, uint8_t* out_rgba8,
                              int32_t size);
/* The Worley detail volume of SKYHIP_MATERIAL_DEFAULT1_WORLEY: level 0 of a size^3 R8 volume,
 * texel (x, y, z) = UNORM8(WorleyFBM(((x, y, z) + 0.5) / size)), the 3D WorleyFBM of NoiseGen.comp:36-50
 * that DETAIL_MAP_GEN (:80-86) mixes into the detail volume, here stored on its own.  No
 * reference counterpart. */
int skyhip_noise_worley(SkyhipContext* ctx, const SkyhipNoiseCreateInfo* worley, uint8_t* out_r8, int32_t size);

/* Cell-corner copies of a REPEAT texture's level 0 for one-load LINEAR fetches (no reference
 * counterpart: a device data layout of the raymarch).  2D RG8, size s a power of two: cell
 * (x, y) holds the texels (x, y), (x + 1, y), (x, y + 1), (x + 1, y + 1) (indices mod s), two
 * bytes each, in that order; cells x fastest, out_cells: s^2 x 8 B.  3D R8, edge s a power of
 * two: cell (x, y, z) holds the texels (x + a, y + b, z + c) mod s in the order a + 2b + 4c;
 * out_cells: s^3 x 8 B.  Both 8-byte aligned. */
int skyhip_texture_cells_2d_rg8(SkyhipContext* ctx, const uint8_t* level0, int32_t size, uint8_t* out_cells);
int skyhip_texture_cells_3d_r8(SkyhipContext* ctx, const uint8_t* level0, int32_t size, uint8_t* out_cells);
/* glGenerateTextureMipmap (DynamicTexture::Generate, VolumetricCloudDefaultMaterial.h:39-49)
 * for a packed UNORM8 chain: fills levels 1..levels-1 from level 0 with a
 * 2x2 (dims = 2) or 2x2x2 (dims = 3) box filter, rounded half up. */
int skyhip_generate_mipmaps(SkyhipContext* ctx, uint8_t* texture, int32_t dims, int32_t channels, int32_t width,
                            int32_t height, int32_t depth, int32_t levels);

/* ------------------------------------------------------------------------- */
/* Volumetric-cloud raymarch                                                  */
/* ------------------------------------------------------------------------- */

/* The render pass of VolumetricCloud::Render (src/SkyRendering/VolumetricCloud.cpp:362-386,
 * shaders/SkyRendering/VolumetricCloudRender.comp:139-210) with the material's
 * SampleSigmaT compiled in.  One ray per output pixel.
 *   depth:          R32F scene depth (GL window z in [0, 1]) at the sample grid, or NULL
 *                   for the far plane everywhere;
 *   shading_index:  NULL for one ray per output pixel (sample grid = output grid), or the
 *                   RG32F index_linear_depth texture of the reference's checkerboard
 *                   pipeline (sample grid = 2x the output grid, VolumetricCloudRender.comp:140-145);
 *   blue_noise:     R16 UNORM tile of blue_noise_size^2 (power of two) read at pixel & (size-1);
 *   out_rgba16f:    RGBA16F (luminance, transmittance); out_distance: R32F mean distance;
 *   work_counters:  NULL, or device uint64[4] to which a counting kernel variant adds the
 *                   executed primary march steps, density evaluations (primary + light march),
 *                   detail-texture fetches and in-cloud steps (sigma >= 1e-5); leave NULL when timing;
 *   pixel_steps:    NULL, or (with work_counters) device uint32[2 * width * height] receiving,
 *                   per pixel, the executed steps and then the scheduled steps
 *                   uint(max(steps * dist / max_dist, 1)), each as segment 0 in bits 0-15 and
 *                   segment 1 in bits 16-31, so a checker can replay the discrete decisions
 *                   (step count, early exit at T < 0.01).
 * The aerial-perspective luminance is multiplied by the ray-scatter visibility of the froxel
 * bound with skyhip_bind_shadow_froxel (VolumetricCloudRender.comp:198), or by 1 when none is
 * bound. */
int skyhip_cloud_raymarch(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                          const SkyhipCloudCommonBufferData* common, const SkyhipCloudBufferData* cloud,
                          const SkyhipCloudMaterial* material, const float* transmittance_rgba,
                          int32_t transmittance_width, int32_t transmittance_height,
                          const float* ap_luminance_rgba, const float* ap_transmittance_rgba, int32_t ap_width,
                          int32_t ap_height, int32_t ap_depth, const uint16_t* blue_noise, int32_t blue_noise_size,
                          const float* depth, const float* shading_index, void* out_rgba16f, float* out_distance,
                          int32_t width, int32_t height, unsigned long long* work_counters,
                          uint32_t* pixel_steps);

/* The material plugin's density, SampleSigmaT(pos, CalHeight01(pos)) (the shader-side
 * contract of IVolumetricCloudMaterial, VolumetricCloudRender.comp:96 and
 * VolumetricCloudCommon.glsl:36-40), at `count` local-frame points (float xyz triples,
 * device memory).  out_sigma_t: device float[count], extinction per km. */
int skyhip_cloud_sample_sigma_t(SkyhipContext* ctx, const SkyhipCloudCommonBufferData* common,
                                const SkyhipCloudMaterial* material, const float* positions_xyz, int32_t count,
                                float* out_sigma_t);

/* ------------------------------------------------------------------------- */
/* Voxel-cloud path tracer                                                    */
/* ------------------------------------------------------------------------- */

/* VolumetricCloud::PathTracing::InitParam (src/SkyRendering/VolumetricCloud.h:158-168) plus
 * the constants the reference bakes into the shader source (VolumetricCloud.cpp:506-519). */
typedef enum SkyhipPrng { SKYHIP_PRNG_WANG_HASH = 0, SKYHIP_PRNG_PCG_HASH = 1 } SkyhipPrng;
typedef enum SkyhipEnvironmentLighting {
    SKYHIP_ENV_OFF = 0,
    SKYHIP_ENV_CONST_ENVIRONMENT_MAP = 1,
    SKYHIP_ENV_GROUND_SINGLE_BOUNCE = 2,
    SKYHIP_ENV_GROUND_MULTI_BOUNCE = 3
} SkyhipEnvironmentLighting;

typedef struct SkyhipPathTracingParameters {
    int32_t max_bounces;            /* kMaxBounces, 128 */
    float region_box_half_width;    /* kCloudHalfWidth, 100 (km) */
    int32_t importance_sampling;    /* IMPORTANCE_SAMPLING, 1 */
    float forward_phase_g;          /* kForwardPhaseG, 0.85 */
    float back_phase_g;             /* kBackPhaseG, -0.15 */
    float forward_scattering_ratio; /* kForwardScatteringRatio, 0.7 */
    int32_t prng;                   /* SkyhipPrng, PCG_HASH */
    int32_t environment_lighting;   /* SkyhipEnvironmentLighting, GROUND_MULTI_BOUNCE */
    float sigma_t_max;              /* kSigmaTMax = material GetSigmaTMax() (1/km) */
    float albedo;                   /* single-scattering albedo; the reference's is 1 */
    float model3[9];                /* kModelMatrix3, column-major: cloud local frame -> atmosphere frame */
} SkyhipPathTracingParameters;

/* Device inputs of the path-trace pass: the LUTs it reads (VolumetricCloud.cpp:546-555)
 * and the optional brick majorants.
 *   sky_view_*:      the sky-view LUT pair (RGBA32F) of the atmosphere renderer whose block is
 *                    passed with the call; the environment term samples it directly, which is
 *                    level 0 of the reference's environment-luminance cube up to resampling
 *                    (AtmosphereRenderer.glsl:248-275);
 *   brick_majorants: NULL to track with the global sigma_t_max as the reference does, or the
 *                    float[ceil(dim/B)^3] output of skyhip_voxel_grid_majorants_sized for a
 *                    VOXEL_F16 material (scaled by the voxel block's uDensity on use);
 *   brick_size:      the brick edge B those majorants were built with: 0 for the default
 *                    (skyhip_voxel_grid_brick_size, 16), else a power of two in [4, 64];
 *   cell_corners:    NULL, or the skyhip_voxel_grid_cells copy of the material's grid: the
 *                    brick-majorant tracker then reads each density lookup's eight texels with
 *                    one 16-byte load instead of eight 2-byte ones.  The same texels and the
 *                    same arithmetic: results are identical with and without it. */
typedef struct SkyhipPathTracingInputs {
    const float* transmittance;
    int32_t transmittance_width;
    int32_t transmittance_height;
    const float* sky_view_luminance;
    const float* sky_view_transmittance;
    int32_t sky_view_width;
    int32_t sky_view_height;
    const float* ap_luminance;
    const float* ap_transmittance;
    int32_t ap_width;
    int32_t ap_height;
    int32_t ap_depth;
    int32_t brick_size;
    const float* brick_majorants;
    const uint16_t* cell_corners;
} SkyhipPathTracingInputs;

/* The synthetic voxel-cloud density of the path-tracing workload (no reference
 * counterpart: the reference loads an OpenVDB file).  Per texel centre p in [0,1]^3:
 *   f       = sum_{o<octaves} 0.5^o PerlinNoise(p, base_frequency 2^o, seed) / sum 0.5^o
 *   falloff = max(1 - |(p - ellipsoid_center) / ellipsoid_radii|^2, 0)
 *   density = max(exp(lognormal_sigma * fbm_inv_std * f) * falloff - threshold, 0)
 * stored as FP16 (round to nearest even); PerlinNoise is Noise.glsl:28-60. */
typedef struct SkyhipVoxelGridCreateInfo {
    uint32_t seed;
    uint32_t base_frequency;
    int32_t octaves;
    float fbm_inv_std;
    float lognormal_sigma;
    float threshold;
    float ellipsoid_center[3];
    float ellipsoid_radii[3];
} SkyhipVoxelGridCreateInfo;

/* Fills out_grid (device FP16[dim_x * dim_y * dim_z], x fastest).  Dims 8 .. 1024. */
int skyhip_voxel_grid_generate(SkyhipContext* ctx, const SkyhipVoxelGridCreateInfo* info, uint16_t* out_grid,
                               int32_t dim_x, int32_t dim_y, int32_t dim_z);

/* Default edge of a majorant brick in texels (16: the fastest edge measured on the 512^3
 * workload grid). */
int skyhip_voxel_grid_brick_size(void);
/* Majorant per B^3 brick of an FP16 grid (B a power of two in [4, 64]): the maximum over the
 * brick's texels and a one-texel apron on every side (clamped to the grid), which bounds
 * trilinear CLAMP_TO_BORDER(0) samples anywhere in the brick.  out_majorants: device
 * float[ceil(dim_x/B) * ceil(dim_y/B) * ceil(dim_z/B)], x fastest.  The reference tracks with
 * one global majorant (VolumetricCloud.h:159-167); local majorants are an addition. */
int skyhip_voxel_grid_majorants_sized(SkyhipContext* ctx, const uint16_t* grid, int32_t dim_x, int32_t dim_y,
                                      int32_t dim_z, int32_t brick_size, float* out_majorants);
/* skyhip_voxel_grid_majorants_sized with the default brick edge. */
int skyhip_voxel_grid_majorants(SkyhipContext* ctx, const uint16_t* grid, int32_t dim_x, int32_t dim_y,
                                int32_t dim_z, float* out_majorants);
/* Cell-corner copy of an FP16 grid, for trilinear lookups with one load (no reference
 * counterpart: a device data layout of the path tracer).  Cell (i, j, k), i in [-1, dim_x - 1]
 * (likewise j, k), holds the eight texels (i + a, j + b, k + c), a, b, c in {0, 1}, in the
 * order a + 2b + 4c, with +0 for texels outside the grid (CLAMP_TO_BORDER(0)); cells are
 * stored x fastest from (-1, -1, -1).  out_cells: device uint16_t[8 * (dim_x + 1) *
 * (dim_y + 1) * (dim_z + 1)], 16-byte aligned (4.2 x the grid: 2.16 GB for 512^3); at most
 * 2^31 cells. */
int skyhip_voxel_grid_cells(SkyhipContext* ctx, const uint16_t* grid, int32_t dim_x, int32_t dim_y, int32_t dim_z,
                            uint16_t* out_cells);

/* One or more frames of VolumetricCloud::PathTracing::Render's trace pass
 * (src/SkyRendering/VolumetricCloud.cpp:537-566, VolumetricCloudPathTracing.comp:255-284):
 * for frame ids first_frame .. first_frame + frames - 1 (the reference's frame_cnt_, from 1)
 * every pixel of region = {x0, y0, x1, y1} traces one path per frame and adds (L, alpha) to
 * accumulation_rgba (RGBA32F width x height, in the frame order), and rendered_mask
 * (R8, may be NULL) gets 1.  The random sequence of each sample is the reference's
 * (seed = PRNG(PRNG(PRNG(x) + y) + frame)); with brick majorants the tracking draws differ
 * but the estimator keeps its expectation.  A scattered path's aerial-perspective luminance is
 * multiplied by the froxel bound with skyhip_bind_shadow_froxel (VolumetricCloudPathTracing.comp:277),
 * or by 1 when none is bound.
 *   work_counters: NULL, or device uint64[5] accumulating samples, density lookups,
 *                  majorant segments, path vertices (volume and ground events) and shadow rays
 *                  (a counting kernel variant; leave NULL when timing).
 * With brick majorants, each call first writes the bricks' inverse majorants (1 / (majorant *
 * uDensity), 0 for an empty brick) into a context workspace of one float per brick, and with
 * bricks of 4 texels and the cell-corner copy bound, each frame is one persistent
 * launch whose waves take 8x8-pixel tiles from a context counter.  The first call that needs
 * them allocates these (the workspace grows to the largest brick grid traced; skyhip_destroy
 * frees both).  Calls on one context run in stream order and share them. */
int skyhip_cloud_path_trace(SkyhipContext* ctx, const SkyhipAtmosphereBufferData* atmosphere,
                            const SkyhipAtmosphereRenderBufferData* render, const SkyhipCloudCommonBufferData* common,
                            const SkyhipCloudMaterial* material, const SkyhipPathTracingParameters* params,
                            const SkyhipPathTracingInputs* inputs, const int32_t region[4], uint32_t first_frame,
                            int32_t frames, float* accumulation_rgba, uint8_t* rendered_mask, int32_t width,
                            int32_t height, unsigned long long* work_counters);

/* The display pass (VolumetricCloudPathTracing.comp:288-296): for every pixel,
 * avg = accumulation / (mask ? frame_id : frame_id - 1) and display.rgb = display.rgb * avg.a + avg.rgb
 * (display_rgba16f: RGBA16F width x height, read and written; rendered_mask may be NULL = all 1). */
int skyhip_cloud_path_trace_display(SkyhipContext* ctx, const float* accumulation_rgba, const uint8_t* rendered_mask,
                                    uint32_t frame_id, void* display_rgba16f, int32_t width, int32_t height);

/* ------------------------------------------------------------------------- */
/* Batched LUT builds (no reference counterpart: the reference builds one     */
/* atmosphere's LUTs per frame, Atmosphere::UpdateLuts and                    */
/* AtmosphereRenderer::Render; these build many at once)                      */
/* ------------------------------------------------------------------------- */

/* One atmosphere set: its block and the sizes of its transmittance and
 * multiple-scattering LUTs (each built exactly as skyhip_transmittance_lut /
 * skyhip_multiscattering_lut build them). */
typedef struct SkyhipLutBatchSet {
    SkyhipAtmosphereBufferData atmosphere;
    int32_t transmittance_width;
    int32_t transmittance_height;
    int32_t multiscattering_width;
    int32_t multiscattering_height;
} SkyhipLutBatchSet;

/* One view of a set (a sun direction and camera): its renderer block and the
 * sky-view LUT size (built exactly as skyhip_sky_view_lut builds it, without dither). */
typedef struct SkyhipLutBatchView {
    int32_t set;
    int32_t sky_view_width;
    int32_t sky_view_height;
    int32_t padding;
    SkyhipAtmosphereRenderBufferData render;
} SkyhipLutBatchView;

/* Launch shape of the three persistent passes: workgroup sizes (64/128/256; multiple
 * scattering: waves of 64 lanes per workgroup, 1/2/4/8), texels per tile, resident
 * workgroups per CU, and whether sky-view tiles of the largest LUTs go first. */
typedef struct SkyhipLutBatchTuning {
    int32_t transmittance_block;
    int32_t transmittance_tile;
    int32_t transmittance_blocks_per_cu;
    int32_t multiscattering_waves;
    int32_t multiscattering_blocks_per_cu;
    int32_t sky_view_block;
    int32_t sky_view_tile;
    int32_t sky_view_blocks_per_cu;
    int32_t sky_view_largest_first;
} SkyhipLutBatchTuning;

typedef struct SkyhipLutBatch SkyhipLutBatch;

/* Uploads the sets and views (host arrays), allocates every output LUT in device pools
 * and captures the build into one hipGraph (default tuning).  Synchronous. */
int skyhip_lut_batch_create(SkyhipContext* ctx, const SkyhipLutBatchSet* sets, int32_t set_count,
                            const SkyhipLutBatchView* views, int32_t view_count, SkyhipLutBatch** out_batch);
int skyhip_lut_batch_destroy(SkyhipLutBatch* batch);
/* The whole build (every set's T and MS, then every view's sky view): one hipGraphLaunch
 * on the context's stream. */
int skyhip_lut_batch_launch(SkyhipLutBatch* batch);
int skyhip_lut_batch_get_tuning(const SkyhipLutBatch* batch, SkyhipLutBatchTuning* out_tuning);
/* Re-captures the graph with the given tuning.  Synchronous. */
int skyhip_lut_batch_set_tuning(SkyhipLutBatch* batch, const SkyhipLutBatchTuning* tuning);
/* Times each pass's candidate launch shapes on this batch (median of `repetitions` runs,
 * two coordinate-descent sweeps), keeps the fastest and re-captures the graph.  out_ms:
 * the sum of the three passes' best times.  Synchronous; leaves valid outputs. */
int skyhip_lut_batch_autotune(SkyhipLutBatch* batch, int32_t repetitions, SkyhipLutBatchTuning* out_best,
                              float* out_ms);
/* Device pointers to a set's / view's output LUTs (RGBA32F, the sizes given at creation). */
int skyhip_lut_batch_set_outputs(SkyhipLutBatch* batch, int32_t set, const float** transmittance_rgba,
                                 const float** multiscattering_rgba);
int skyhip_lut_batch_view_outputs(SkyhipLutBatch* batch, int32_t view, const float** luminance_rgba,
                                  const float** sky_transmittance_rgba);
/* Integrator steps of one launch: T texels x transmittance loop iterations + MS texels x 64
 * directions x MS iterations + sky-view texels x sky-view iterations. */
int skyhip_lut_batch_integrator_steps(const SkyhipLutBatch* batch, unsigned long long* out_steps);

/* ------------------------------------------------------------------------- */
/* Learned sky-view surrogate (no reference counterpart: an addition that     */
/* replaces the per-pixel sky raymarch of AtmosphereRenderer.glsl:372-379     */
/* with a small MLP, checked at run time against that raymarch)               */
/* ------------------------------------------------------------------------- */

/* A sky sample: view from `altitude` (km above the ground) along a direction with
 * zenith cosine mu_view, sun zenith cosine mu_sun and azimuth difference delta_phi
 * in [0, pi], in atmosphere `atmosphere` of a SkyhipSkyLuts set. */
typedef struct SkyhipSkyQuery {
    int32_t atmosphere;
    float altitude;
    float mu_view;
    float mu_sun;
    float delta_phi;
    float pad[3];
} SkyhipSkyQuery;

/* A panorama view: camera altitude and sun zenith cosine in one atmosphere. */
typedef struct SkyhipSkyView {
    int32_t atmosphere;
    float altitude;
    float mu_sun;
    float pad;
} SkyhipSkyView;

/* Device LUTs of several atmospheres with uniform sizes: atmosphere a's blocks at
 * atmospheres[a], its transmittance LUT at transmittance + a * tw * th * 4 floats and its
 * multiple-scattering LUT at multiscattering + a * mw * mh * 4 floats (the pool layout of a
 * LUT batch whose sets share one size).  steps: integrator steps per ray. */
typedef struct SkyhipSkyLuts {
    const SkyhipAtmosphereBufferData* atmospheres;
    const float* transmittance;
    const float* multiscattering;
    int32_t transmittance_width;
    int32_t transmittance_height;
    int32_t multiscattering_width;
    int32_t multiscattering_height;
    float steps;
    int32_t atmosphere_count; /* query / view atmosphere indices are clamped to [0, count - 1] */
} SkyhipSkyLuts;

/* 13187: W1 (64x7), b1, W2..W4 (64x64), b2..b4, W5 (3x64), b5, each W out-major. */
int skyhip_surrogate_parameter_count(void);
/* The network kind, an argument of the forward, gradient and render entry points: the same
 * layers and parameters with
 *  - SKYHIP_SURROGATE_F16_GATE: FP16 MFMA operands and the gate a(z) = z t^2 (3 - 2 t),
 *    t = clamp(z / 6 + 1/2, 0, 1), evaluated in packed FP16 (the default), or
 *  - SKYHIP_SURROGATE_BF16_SILU: BF16 MFMA operands and SiLU z sigmoid(z) in FP32.
 * Both accumulate in FP32. */
enum {
    SKYHIP_SURROGATE_F16_GATE = 0,
    SKYHIP_SURROGATE_BF16_SILU = 1
};
/* The plain integrator per query (ComputeScatteredLuminance with the multiple-scattering
 * term, Atmosphere.glsl:220-295, along the query's ray): out_rgb device float[3 * count]. */
int skyhip_sky_radiance(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyQuery* queries, int32_t count,
                        float* out_rgb);
/* Surrogate inputs (8 floats per query, the 8th 0): 2 sqrt(altitude / thickness) - 1; the
 * view zenith as the sky-view LUT's latitude coordinate x_lat (square-root spacing from the
 * horizon on both sides, GetSkyViewTextureUvFromCosLatLon), 2 x_lat - 1;
 * sign(mu_sun) sqrt(|mu_sun|); the
 * sun-view cosine mu_view mu_sun + sin_view sin_sun cos(delta_phi); and the atmosphere's
 * natural-log Rayleigh, Mie and ozone scale factors (log_scales: device
 * float[3 * atmospheres]) divided by ln 4. */
int skyhip_surrogate_features(SkyhipContext* ctx, const SkyhipSkyQuery* queries, int32_t count,
                              const float* log_scales, int32_t atmosphere_count, float thickness,
                              float* out_features);
/* Inference (MFMA operands and activation per `network`, FP32 accumulate): out device
 * float[3 * count] log radiance. */
int skyhip_surrogate_forward(SkyhipContext* ctx, int32_t network, const float* params, const float* features,
                             int32_t count, float* out_log_radiance);
/* Workspace of skyhip_surrogate_gradient for `count` samples. */
size_t skyhip_surrogate_workspace_bytes(int32_t count);
/* Gradient of loss = loss_scale * sum_samples w * sum_rgb (f(x) - target)^2 over `count`
 * samples (targets: device float[3 * count] log radiance; weights: device float[count], or
 * NULL for w = 1) by the MFMA forward and backward kernels of `network`; out_grad: device
 * float[13187]; out_loss: device float (may be NULL). */
int skyhip_surrogate_gradient(SkyhipContext* ctx, int32_t network, const float* params, const float* features,
                              const float* targets, const float* weights, int32_t count, float loss_scale,
                              float* out_grad, float* out_loss, void* workspace, size_t workspace_bytes);
/* skyhip_surrogate_gradient without a workspace argument and without allocating: for callers that
 * cannot allocate (code under HIP graph capture, tight training loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  The partial gradients live in a fixed device array, one
 * per GPU, resolved when a context is created on that GPU and held until the process ends
 * (13.5 MB); count must be at most SKYHIP_SURROGATE_NOALLOC_MAX_COUNT (larger counts fail with
 * SKYHIP_ERR_INVALID_SIZE; use skyhip_surrogate_gradient).  All calls on one GPU share the array:
 * calls that overlap on one GPU (on different streams or from different threads) race. */
#define SKYHIP_SURROGATE_NOALLOC_MAX_COUNT 8192
int skyhip_surrogate_gradient_noalloc(SkyhipContext* ctx, int32_t network, const float* params,
                                      const float* features, const float* targets, const float* weights,
                                      int32_t count, float loss_scale, float* out_grad, float* out_loss);
/* One Adam step (step counts from 1) on device params / moments. */
int skyhip_surrogate_adam(SkyhipContext* ctx, float* params, const float* grad, float* m, float* v, float lr,
                          float beta1, float beta2, float eps, int32_t step);
/* Equirectangular panoramas (width x height per view; pixel (x, y): zenith angle
 * pi (y + 0.5) / height, azimuth from the sun 2 pi (x + 0.5) / width - pi) of RGB sky
 * radiance (out_rgba device float[4 * width * height * view_count]) by the surrogate,
 * with a residual check per 16 x 16 tile: the plain integrator at four probe pixels, and
 * if any probe's max_c |surrogate - plain| / max(plain, radiance_floor) exceeds
 * `tolerance` the whole tile is recomputed by the plain integrator.  tile_flags: device
 * uint32[view_count * ceil(width/16) * ceil(height/16)] (1 = fell back); counters: device
 * uint32[1], the number of tiles that fell back.  network: as for skyhip_surrogate_forward. */
int skyhip_sky_render_surrogate(SkyhipContext* ctx, int32_t network, const float* params, const SkyhipSkyLuts* luts,
                                const SkyhipSkyView* views, int32_t view_count, const float* log_scales,
                                float thickness, int32_t width, int32_t height, float tolerance,
                                float radiance_floor, float* out_rgba, uint32_t* tile_flags, uint32_t* counters);
/* The same panoramas by the plain integrator at every pixel. */
int skyhip_sky_render_plain(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyView* views,
                            int32_t view_count, int32_t width, int32_t height, float* out_rgba);
/* The same panoramas by the reference's method without the surrogate: each view's sky-view
 * LUT (sky_view_luminance: device array of view_count device pointers to RGBA32F LUTs of
 * sky_view_width x sky_view_height, built for the view's altitude and sun) sampled
 * bilinearly at the pixel's direction (AtmosphereRenderer.glsl:113-145, 372-379). */
int skyhip_sky_render_lut(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyView* views,
                          int32_t view_count, const float* const* sky_view_luminance, int32_t sky_view_width,
                          int32_t sky_view_height, int32_t width, int32_t height, float* out_rgba);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* SKYHIP_H */
