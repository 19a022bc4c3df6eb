// Host side of the skyhip C ABI: contexts, device memory, timing events,
// error strings, and the reference's host-side parameter conversions
// (AssignBufferData / Camera / VolumetricCloud::Update), restated in C++ with
// the reference's float arithmetic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstring>

#include "common.h"

namespace skyhip {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return SKYHIP_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? SKYHIP_ERR_OUT_OF_MEMORY : SKYHIP_ERR_DEVICE;
}

bool finite_block(const float* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

int require(bool cond, int code, const char* what) {
    if (cond) return SKYHIP_OK;
    set_error("%s", what);
    return code;
}

int check_ctx(SkyhipContext* ctx) {
    if (ctx == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "context is NULL");
    if (ctx->stream == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "context has no stream");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return check_hip(e, "hipSetDevice");
    return SKYHIP_OK;
}

// ---------------------------------------------------------------------------
// Small float vector / matrix helpers with glm semantics (column-major mat4).
// ---------------------------------------------------------------------------
struct V3f {
    float x, y, z;
};
static V3f operator+(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static V3f operator-(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static V3f operator*(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static float dot3(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3f cross3(V3f a, V3f b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
static float length3(V3f a) { return std::sqrt(dot3(a, a)); }
static V3f normalize3(V3f a) { return a * (1.0f / std::sqrt(dot3(a, a))); }
static V3f load3(const float* p) { return {p[0], p[1], p[2]}; }
static void store3(float* p, V3f v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

// Utils.cpp:23-31
static V3f from_theta_phi(float theta, float phi) {
    return {std::cos(phi) * std::sin(theta), std::cos(theta), std::sin(phi) * std::sin(theta)};
}

static void mat_identity(float* m) {
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}
// out = a * b (column-major: element (row r, col c) at [c * 4 + r])
static void mat_mul(const float* a, const float* b, float* out) {
    float t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            t[c * 4 + r] = s;
        }
    std::memcpy(out, t, sizeof(t));
}
// General 4x4 inverse by cofactors, evaluated in double and rounded once.
static bool mat_inverse(const float* m, float* out) {
    double a[16], inv[16];
    for (int i = 0; i < 16; ++i) a[i] = m[i];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
             a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
             a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
             a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
              a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
             a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
             a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
             a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
              a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
             a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
             a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
              a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
              a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
             a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
             a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
              a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
              a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) out[i] = float(inv[i] / det);
    return true;
}

// glm::lookAt (right-handed)
static void look_at(V3f eye, V3f center, V3f up, float* m) {
    V3f f = normalize3(center - eye);
    V3f s = normalize3(cross3(f, up));
    V3f u = cross3(s, f);
    mat_identity(m);
    m[0] = s.x; m[4] = s.y; m[8] = s.z;
    m[1] = u.x; m[5] = u.y; m[9] = u.z;
    m[2] = -f.x; m[6] = -f.y; m[10] = -f.z;
    m[12] = -dot3(s, eye);
    m[13] = -dot3(u, eye);
    m[14] = dot3(f, eye);
}

// glm::ortho(left, right, bottom, top, z_near, z_far): right-handed, clip z in [-1, 1].
static void ortho(float l, float r, float b, float t, float n, float f, float* m) {
    mat_identity(m);
    m[0] = 2.0f / (r - l);
    m[5] = 2.0f / (t - b);
    m[10] = -2.0f / (f - n);
    m[12] = -(r + l) / (r - l);
    m[13] = -(t + b) / (t - b);
    m[14] = -(f + n) / (f - n);
}

// glm::perspective (right-handed, clip z in [-1, 1])
static void perspective(float fovy, float aspect, float z_near, float z_far, float* m) {
    float t = std::tan(fovy / 2.0f);
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = 1.0f / (aspect * t);
    m[5] = 1.0f / t;
    m[10] = -(z_far + z_near) / (z_far - z_near);
    m[11] = -1.0f;
    m[14] = -(2.0f * z_far * z_near) / (z_far - z_near);
}

}  // namespace skyhip

using namespace skyhip;

// ---------------------------------------------------------------------------
// Context / memory / events
// ---------------------------------------------------------------------------
extern "C" int skyhip_create(int device, SkyhipContext** out_ctx) {
    if (out_ctx == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "out_ctx is NULL");
    *out_ctx = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(SKYHIP_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= count)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "device %d outside [0, %d)", device, count);
    SKYHIP_HIP(hipSetDevice(device));
    auto* ctx = new SkyhipContext();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->cu_count = prop.multiProcessorCount;
        ctx->wave_size = prop.warpSize;
        ctx->clock_khz = prop.clockRate;
        std::snprintf(ctx->name, sizeof(ctx->name), "%s", prop.gcnArchName);
    }
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return check_hip(e, "hipStreamCreateWithFlags");
    }
    e = hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_event, hipEventDisableTiming);
    if (e != hipSuccess) {
        skyhip_destroy(ctx);
        return check_hip(e, "auxiliary stream / events");
    }
    for (auto& ev : ctx->events) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            skyhip_destroy(ctx);
            return check_hip(e, "hipEventCreate");
        }
    }
    // This device's copy of the allocation-free gradient's workspace (with the device current).
    const int rc = skyhip::surrogate_scratch_setup(device);
    if (rc != SKYHIP_OK) {
        skyhip_destroy(ctx);
        return rc;
    }
    *out_ctx = ctx;
    return SKYHIP_OK;
}

extern "C" int skyhip_destroy(SkyhipContext* ctx) {
    if (ctx == nullptr) return SKYHIP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux_stream) (void)hipStreamSynchronize(ctx->aux_stream);
    for (auto& ev : ctx->events)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->fork_event) (void)hipEventDestroy(ctx->fork_event);
    if (ctx->join_event) (void)hipEventDestroy(ctx->join_event);
    if (ctx->pt_queue) (void)hipFree(ctx->pt_queue);
    if (ctx->pt_inv_majorants) (void)hipFree(ctx->pt_inv_majorants);
    if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SKYHIP_OK;
}

extern "C" const char* skyhip_last_error(void) { return g_last_error; }

extern "C" void* skyhip_stream(SkyhipContext* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int skyhip_bind_shadow_froxel(SkyhipContext* ctx, const uint16_t* froxel_r16, int32_t width,
                                         int32_t height, int32_t depth) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (froxel_r16 && (width < 1 || height < 1 || depth < 1 || width > 8192 || height > 8192 || depth > 1024))
        return fail(SKYHIP_ERR_INVALID_SIZE, "froxel size out of range");
    ctx->froxel = froxel_r16;
    ctx->froxel_w = froxel_r16 ? width : 0;
    ctx->froxel_h = froxel_r16 ? height : 0;
    ctx->froxel_d = froxel_r16 ? depth : 0;
    return SKYHIP_OK;
}

extern "C" int skyhip_bind_cloud_phase(SkyhipContext* ctx, const SkyhipCloudPhase* phase) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (phase == nullptr) {
        ctx->phase = SkyhipCloudPhase{-0.15f, 2.16f, 0.85f, 0.0f};  // VolumetricCloudRender.comp:125-126
        return SKYHIP_OK;
    }
    if (!(std::fabs(phase->back_g) < 1.0f && std::fabs(phase->forward_g) < 1.0f && std::isfinite(phase->back_scale)))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "phase lobes need |g| < 1 and a finite back_scale");
    ctx->phase = *phase;
    ctx->phase.padding = 0.0f;
    return SKYHIP_OK;
}

extern "C" int skyhip_synchronize(SkyhipContext* ctx) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_HIP(hipStreamSynchronize(ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_malloc(SkyhipContext* ctx, size_t bytes, void** out) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (out == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "out pointer is NULL");
    *out = nullptr;
    if (bytes == 0) return fail(SKYHIP_ERR_INVALID_SIZE, "zero-byte allocation");
    SKYHIP_HIP(hipMalloc(out, bytes));
    return SKYHIP_OK;
}

extern "C" int skyhip_free(SkyhipContext* ctx, void* p) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (p == nullptr) return SKYHIP_OK;
    SKYHIP_HIP(hipStreamSynchronize(ctx->stream));
    SKYHIP_HIP(hipFree(p));
    return SKYHIP_OK;
}

extern "C" int skyhip_memcpy_h2d(SkyhipContext* ctx, void* dst, const void* src, size_t bytes) {
    SKYHIP_CHECK(check_ctx(ctx));
    if ((dst == nullptr || src == nullptr) && bytes) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL copy operand");
    if (bytes == 0) return SKYHIP_OK;
    SKYHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SKYHIP_HIP(hipStreamSynchronize(ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_memcpy_d2h(SkyhipContext* ctx, void* dst, const void* src, size_t bytes) {
    SKYHIP_CHECK(check_ctx(ctx));
    if ((dst == nullptr || src == nullptr) && bytes) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL copy operand");
    if (bytes == 0) return SKYHIP_OK;
    SKYHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SKYHIP_HIP(hipStreamSynchronize(ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_memcpy_d2d(SkyhipContext* ctx, void* dst, const void* src, size_t bytes) {
    SKYHIP_CHECK(check_ctx(ctx));
    if ((dst == nullptr || src == nullptr) && bytes) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL copy operand");
    if (bytes == 0) return SKYHIP_OK;
    SKYHIP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_memset(SkyhipContext* ctx, void* dst, int value, size_t bytes) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (dst == nullptr && bytes) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL memset target");
    if (bytes == 0) return SKYHIP_OK;
    SKYHIP_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_event_record(SkyhipContext* ctx, int slot) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (slot < 0 || slot >= 64) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "event slot %d outside [0, 64)", slot);
    SKYHIP_HIP(hipEventRecord(ctx->events[slot], ctx->stream));
    return SKYHIP_OK;
}

extern "C" int skyhip_event_elapsed_ms(SkyhipContext* ctx, int start_slot, int end_slot, float* out_ms) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (start_slot < 0 || start_slot >= 64 || end_slot < 0 || end_slot >= 64 || out_ms == nullptr)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "bad event slots or NULL output");
    SKYHIP_HIP(hipEventSynchronize(ctx->events[end_slot]));
    SKYHIP_HIP(hipEventElapsedTime(out_ms, ctx->events[start_slot], ctx->events[end_slot]));
    return SKYHIP_OK;
}

extern "C" int skyhip_device_info(SkyhipContext* ctx, int32_t* cu_count, int32_t* wave_size, int32_t* clock_khz,
                                  char* name, int32_t name_len) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (cu_count) *cu_count = ctx->cu_count;
    if (wave_size) *wave_size = ctx->wave_size;
    if (clock_khz) *clock_khz = ctx->clock_khz;
    if (name && name_len > 0) std::snprintf(name, size_t(name_len), "%s", ctx->name);
    return SKYHIP_OK;
}

// ---------------------------------------------------------------------------
// Host-side parameter conversion
// ---------------------------------------------------------------------------
extern "C" int skyhip_atmosphere_assign_buffer_data(const SkyhipAtmosphereParameters* p,
                                                    SkyhipAtmosphereBufferData* d) {
    if (p == nullptr || d == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL parameters or output");
    if (!finite_block(reinterpret_cast<const float*>(p), sizeof(*p) / 4))
        return fail(SKYHIP_ERR_NON_FINITE, "atmosphere parameters have a non-finite field");
    if (p->rayleigh_exponential_distribution == 0.0f || p->mie_exponential_distribution == 0.0f ||
        p->ozone_width == 0.0f)
        return fail(SKYHIP_ERR_NON_FINITE, "zero scale height or ozone width (its inverse is infinite)");
    std::memset(d, 0, sizeof(*d));
    // Atmosphere.cpp:49-70
    store3(d->solar_illuminance, load3(p->solar_illuminance));
    d->sun_angular_radius = radians(p->sun_angular_radius);
    store3(d->rayleigh_scattering, load3(p->rayleigh_scattering) * p->rayleigh_scattering_scale);
    store3(d->mie_scattering, load3(p->mie_scattering) * p->mie_scattering_scale);
    store3(d->mie_absorption, load3(p->mie_absorption) * p->mie_absorption_scale);
    store3(d->ozone_absorption, load3(p->ozone_absorption) * p->ozone_absorption_scale);
    d->inv_rayleigh_exponential_distribution = 1.0f / p->rayleigh_exponential_distribution;
    d->inv_mie_exponential_distribution = 1.0f / p->mie_exponential_distribution;
    d->ozone_center_altitude = p->ozone_center_altitude;
    d->inv_ozone_width = 1.0f / p->ozone_width;
    store3(d->ground_albedo, load3(p->ground_albedo));
    d->bottom_radius = p->bottom_radius;
    d->top_radius = p->bottom_radius + p->thickness;
    d->mie_phase_g = p->mie_phase_g;
    d->transmittance_steps = p->transmittance_steps;
    d->multiscattering_steps = p->multiscattering_steps;
    d->multiscattering_mask = p->multiscattering_mask;
    return SKYHIP_OK;
}

extern "C" int skyhip_atmosphere_render_assign_buffer_data(const SkyhipAtmosphereRenderParameters* p,
                                                           float bottom_radius,
                                                           SkyhipAtmosphereRenderBufferData* d) {
    if (p == nullptr || d == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL parameters or output");
    if (!finite_block(reinterpret_cast<const float*>(p), sizeof(*p) / 4) || !std::isfinite(bottom_radius))
        return fail(SKYHIP_ERR_NON_FINITE, "render parameters have a non-finite field");
    std::memset(d, 0, sizeof(*d));
    // AtmosphereRenderer.cpp:52-83
    V3f sun = from_theta_phi(radians(p->sun_direction_theta), radians(p->sun_direction_phi));
    store3(d->sun_direction, sun);
    d->star_luminance_scale = p->star_luminance_scale;
    V3f cam = load3(p->camera_position);
    store3(d->camera_position, cam);
    if (!mat_inverse(p->view_projection, d->inv_view_projection))
        return fail(SKYHIP_ERR_NON_FINITE, "view_projection is singular");
    mat_identity(d->light_view_projection);
    mat_identity(d->uCloudShadowMapMat);
    d->raymarching_steps = p->raymarching_steps;
    d->sky_view_lut_steps = p->sky_view_lut_steps;
    d->aerial_perspective_lut_steps = p->aerial_perspective_lut_steps;
    d->aerial_perspective_lut_max_distance = p->aerial_perspective_lut_max_distance;
    V3f center = {0.0f, -bottom_radius, 0.0f};  // Earth::center(), Earth.h:37
    store3(d->earth_center, center);
    d->camera_earth_center_distance = length3(cam - center);
    V3f up = normalize3(cam - center);
    store3(d->up_direction, up);
    V3f right = cross3(sun, up);
    float right_length = length3(right);
    if (right_length != 0.0f) {
        right = right * (1.0f / right_length);
    } else {
        V3f biased = from_theta_phi(radians(p->sun_direction_theta + 90.0f), radians(p->sun_direction_phi));
        right = normalize3(cross3(biased, up));
    }
    store3(d->right_direction, right);
    store3(d->front_direction, cross3(up, right));
    return SKYHIP_OK;
}

extern "C" int skyhip_camera_view_projection(const SkyhipCamera* c, float out_vp[16], float out_front[3],
                                             float out_right[3], float out_up[3]) {
    if (c == nullptr || out_vp == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL camera or output");
    if (!finite_block(reinterpret_cast<const float*>(c), sizeof(*c) / 4))
        return fail(SKYHIP_ERR_NON_FINITE, "camera has a non-finite field");
    if (!(c->z_near > 0.0f) || !(c->z_far > c->z_near) || !(c->aspect > 0.0f) || !(c->fovy > 0.0f && c->fovy < 180.0f))
        return fail(SKYHIP_ERR_NON_FINITE, "degenerate camera projection");
    // Camera::UpdateVectors (Camera.cpp:33-37), kWorldUp = (0, 1, 0)
    V3f world_up = {0.0f, 1.0f, 0.0f};
    V3f front = from_theta_phi(radians(90.0f - c->pitch), radians(c->yaw));
    V3f right = normalize3(cross3(front, world_up));
    V3f up = normalize3(cross3(right, front));
    V3f pos = load3(c->position);
    float view[16], proj[16];
    look_at(pos, pos + front, up, view);                                          // Camera.cpp:13-15
    perspective(radians(c->fovy), c->aspect, c->z_near, c->z_far, proj);          // Camera.cpp:17-19
    mat_mul(proj, view, out_vp);                                                  // Camera.h:20
    if (out_front) store3(out_front, front);
    if (out_right) store3(out_right, right);
    if (out_up) store3(out_up, up);
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_assign_buffer_data(const SkyhipVolumetricCloudParameters* cloud,
                                               const SkyhipCamera* camera, float bottom_radius,
                                               const float sun_direction_world[3],
                                               float aerial_perspective_lut_max_distance, int32_t frame_id,
                                               SkyhipCloudCommonBufferData* out_common,
                                               SkyhipCloudBufferData* out_cloud, float out_model[16]) {
    if (!cloud || !camera || !sun_direction_world || !out_common || !out_cloud)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!finite_block(reinterpret_cast<const float*>(cloud), sizeof(*cloud) / 4) ||
        !finite_block(sun_direction_world, 3) || !std::isfinite(bottom_radius) ||
        !std::isfinite(aerial_perspective_lut_max_distance))
        return fail(SKYHIP_ERR_NON_FINITE, "cloud parameters have a non-finite field");
    if (!(cloud->shadow_froxel_max_distance != 0.0f))
        return fail(SKYHIP_ERR_NON_FINITE, "shadow_froxel_max_distance is zero");
    float vp[16];
    SKYHIP_CHECK(skyhip_camera_view_projection(camera, vp, nullptr, nullptr, nullptr));
    // VolumetricCloud.cpp:188-202
    V3f earth_center = {0.0f, -bottom_radius, 0.0f};
    V3f camera_pos = load3(camera->position);
    V3f up = normalize3(camera_pos - earth_center);
    V3f origin = earth_center + up * bottom_radius;
    V3f front = normalize3(cross3(up, V3f{1.0f, 0.0f, 0.0f}));
    V3f right = normalize3(cross3(up, front));
    float model[16] = {front.x, front.y, front.z, 0.0f, right.x, right.y, right.z, 0.0f,
                       up.x,    up.y,    up.z,    0.0f, origin.x, origin.y, origin.z, 1.0f};
    float inv_model[16], mvp[16];
    if (!mat_inverse(model, inv_model)) return fail(SKYHIP_ERR_NON_FINITE, "degenerate local cloud frame");
    mat_mul(vp, model, mvp);
    // :219-220
    V3f local_cam = {inv_model[0] * camera_pos.x + inv_model[4] * camera_pos.y + inv_model[8] * camera_pos.z + inv_model[12],
                     inv_model[1] * camera_pos.x + inv_model[5] * camera_pos.y + inv_model[9] * camera_pos.z + inv_model[13],
                     inv_model[2] * camera_pos.x + inv_model[6] * camera_pos.y + inv_model[10] * camera_pos.z + inv_model[14]};
    V3f sw = load3(sun_direction_world);
    V3f local_sun = normalize3(V3f{inv_model[0] * sw.x + inv_model[4] * sw.y + inv_model[8] * sw.z,
                                   inv_model[1] * sw.x + inv_model[5] * sw.y + inv_model[9] * sw.z,
                                   inv_model[2] * sw.x + inv_model[6] * sw.y + inv_model[10] * sw.z});
    SkyhipCloudCommonBufferData& c = *out_common;
    std::memset(&c, 0, sizeof(c));
    if (!mat_inverse(mvp, c.uInvMVP)) return fail(SKYHIP_ERR_NON_FINITE, "MVP is singular");
    mat_identity(c.uReprojectMat);
    mat_identity(c.uLightVP);
    mat_identity(c.uInvLightVP);
    mat_identity(c.uShadowMapReprojectMat);
    store3(c.uCameraPos, local_cam);
    c.uBaseShadingIndex = uint32_t(frame_id) & 0x3u;
    c.uLinearDepthParam[0] = 1.0f / camera->z_near;
    c.uLinearDepthParam[1] = (camera->z_far - camera->z_near) / (camera->z_far * camera->z_near);
    c.uEarthRadius = bottom_radius;
    store3(c.uSunDirection, local_sun);
    c.uBottomAltitude = cloud->bottom_altitude;
    c.uTopAltitude = cloud->bottom_altitude + cloud->thickness;
    c.uFrameID = float(frame_id);
    c.uShadowFroxelMaxDistance = cloud->shadow_froxel_max_distance;
    c.uAerialPerspectiveLutMaxDistance = aerial_perspective_lut_max_distance;
    c.uInvShadowFroxelMaxDistance = 1.0f / cloud->shadow_froxel_max_distance;
    // :258-269
    SkyhipCloudBufferData& b = *out_cloud;
    std::memset(&b, 0, sizeof(b));
    b.uMaxRaymarchDistance = cloud->max_raymarch_distance;
    b.uMaxRaymarchSteps = cloud->max_raymarch_steps;
    b.uMaxVisibleDistance = cloud->max_visible_distance;
    store3(b.uEnvColorScale, load3(cloud->env_color) * cloud->env_color_scale);
    b.uSunIlluminanceScale = cloud->sun_illuminance_scale;
    b.uShadowSteps = cloud->shadow_steps;
    b.uShadowDistance = cloud->shadow_distance;
    b.uSunMultiscatteringSigmaScale = cloud->sun_multiscattering_sigma_scale;
    b.uEnvMultiscatteringSigmaScale = cloud->env_multiscattering_sigma_scale;
    b.uEnvBottomVisibility = cloud->env_bottom_visibility;
    b.uEnvSunHeightCurveExp = cloud->env_sun_height_curve_exp;
    if (out_model) std::memcpy(out_model, model, sizeof(model));
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_shadow_matrices(const SkyhipVolumetricCloudParameters* cloud, const SkyhipCamera* camera,
                                            float bottom_radius, const float sun_direction_world[3],
                                            int32_t shadow_map_width, int32_t shadow_map_height,
                                            const float previous_light_vp[16], const float previous_model[16],
                                            const float additional_delta[2], SkyhipCloudCommonBufferData* common,
                                            float out_light_vp[16], float out_light_vp_inv_model[16]) {
    if (!cloud || !camera || !sun_direction_world || !common) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((previous_light_vp == nullptr) != (previous_model == nullptr))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "previous_light_vp and previous_model go together");
    if (shadow_map_width < 1 || shadow_map_height < 1) return fail(SKYHIP_ERR_INVALID_SIZE, "shadow map size");
    if (!finite_block(sun_direction_world, 3) || !std::isfinite(bottom_radius) ||
        !std::isfinite(cloud->shadow_map_max_distance) || !(cloud->shadow_map_max_distance > 0.0f))
        return fail(SKYHIP_ERR_NON_FINITE, "non-finite sun direction, radius or shadow distance");
    float vp[16];
    V3f front, right, up_c;
    {
        float f3_[3], r3_[3], u3_[3];
        SKYHIP_CHECK(skyhip_camera_view_projection(camera, vp, f3_, r3_, u3_));
        front = load3(f3_);
        right = load3(r3_);
        up_c = load3(u3_);
    }
    // The local frame of VolumetricCloud::Update (VolumetricCloud.cpp:188-199), as in
    // skyhip_cloud_assign_buffer_data.
    const V3f earth_center = {0.0f, -bottom_radius, 0.0f};
    const V3f camera_pos = load3(camera->position);
    const V3f up = normalize3(camera_pos - earth_center);
    const V3f origin = earth_center + up * bottom_radius;
    const V3f fr = normalize3(cross3(up, V3f{1.0f, 0.0f, 0.0f}));
    const V3f ri = normalize3(cross3(up, fr));
    const float model[16] = {fr.x, fr.y, fr.z, 0.0f, ri.x, ri.y, ri.z, 0.0f,
                             up.x, up.y, up.z, 0.0f, origin.x, origin.y, origin.z, 1.0f};
    float inv_model[16];
    if (!mat_inverse(model, inv_model)) return fail(SKYHIP_ERR_NON_FINITE, "degenerate local cloud frame");
    auto to_local = [&](V3f p, float w) {
        return V3f{inv_model[0] * p.x + inv_model[4] * p.y + inv_model[8] * p.z + inv_model[12] * w,
                   inv_model[1] * p.x + inv_model[5] * p.y + inv_model[9] * p.z + inv_model[13] * w,
                   inv_model[2] * p.x + inv_model[6] * p.y + inv_model[10] * p.z + inv_model[14] * w};
    };
    const V3f local_cam = to_local(camera_pos, 1.0f);
    const V3f local_sun = normalize3(to_local(load3(sun_direction_world), 0.0f));
    // :221-223: the light view looks from the camera along -sun.
    const V3f light_up = (local_sun.x == 0.0f && local_sun.y == 0.0f) ? V3f{1.0f, 0.0f, 0.0f} : V3f{0.0f, 0.0f, 1.0f};
    float light_view[16];
    look_at(local_cam, local_cam - local_sun, light_up, light_view);
    // GetLightProjection (:152-172): the light-space bounds of the camera and the far-plane
    // corners at shadow_map_max_distance, padded by half a texel, as an orthographic box.
    const float dmax = cloud->shadow_map_max_distance;
    const float tan_half = std::tan(radians(camera->fovy) * 0.5f);
    const V3f far_center = camera_pos + front * dmax;
    const V3f u = up_c * (tan_half * dmax), rr = right * (camera->aspect * tan_half * dmax);
    const V3f corners[5] = {camera_pos, far_center + u + rr, far_center + u - rr, far_center - u + rr,
                            far_center - u - rr};
    float min_xy[2] = {1e10f, 1e10f}, max_xy[2] = {-1e10f, -1e10f};
    for (const V3f& cw : corners) {
        const V3f l = to_local(cw, 1.0f);
        const float lx = light_view[0] * l.x + light_view[4] * l.y + light_view[8] * l.z + light_view[12];
        const float ly = light_view[1] * l.x + light_view[5] * l.y + light_view[9] * l.z + light_view[13];
        min_xy[0] = std::min(min_xy[0], lx);
        max_xy[0] = std::max(max_xy[0], lx);
        min_xy[1] = std::min(min_xy[1], ly);
        max_xy[1] = std::max(max_xy[1], ly);
    }
    const float pad_x = 0.5f * (max_xy[0] - min_xy[0]) / float(shadow_map_width);
    const float pad_y = 0.5f * (max_xy[1] - min_xy[1]) / float(shadow_map_height);
    float light_proj[16], light_vp[16], inv_light_vp[16];
    ortho(min_xy[0] - pad_x, max_xy[0] + pad_x, min_xy[1] - pad_y, max_xy[1] + pad_y, -1.0f, 1.0f, light_proj);
    mat_mul(light_proj, light_view, light_vp);
    if (!mat_inverse(light_vp, inv_light_vp)) return fail(SKYHIP_ERR_NON_FINITE, "light view-projection is singular");
    std::memcpy(common->uLightVP, light_vp, sizeof(light_vp));
    std::memcpy(common->uInvLightVP, inv_light_vp, sizeof(inv_light_vp));
    // :235: previous light VP * (previous model^-1 * model) * additional-delta translation * light VP^-1;
    // identity without a previous frame.
    if (previous_light_vp) {
        float inv_prev_model[16], rel[16], delta[16], t0[16], t1[16];
        if (!mat_inverse(previous_model, inv_prev_model)) return fail(SKYHIP_ERR_NON_FINITE, "previous model singular");
        mat_mul(inv_prev_model, model, rel);
        mat_identity(delta);
        if (additional_delta) {
            delta[12] = additional_delta[0];
            delta[13] = additional_delta[1];
        }
        mat_mul(previous_light_vp, rel, t0);
        mat_mul(t0, delta, t1);
        mat_mul(t1, inv_light_vp, common->uShadowMapReprojectMat);
    } else {
        mat_identity(common->uShadowMapReprojectMat);
    }
    if (out_light_vp) std::memcpy(out_light_vp, light_vp, sizeof(light_vp));
    if (out_light_vp_inv_model) mat_mul(light_vp, inv_model, out_light_vp_inv_model);  // :277
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_reproject_matrix(const SkyhipCamera* camera, float bottom_radius,
                                             const float previous_mvp[16], const float previous_camera_position[3],
                                             const float additional_delta[2], SkyhipCloudCommonBufferData* common,
                                             float out_mvp[16]) {
    if (!camera || !common) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((previous_mvp == nullptr) != (previous_camera_position == nullptr))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "previous_mvp and previous_camera_position go together");
    if ((previous_mvp && (!finite_block(previous_mvp, 16) || !finite_block(previous_camera_position, 3))) ||
        (additional_delta && !finite_block(additional_delta, 2)) || !std::isfinite(bottom_radius))
        return fail(SKYHIP_ERR_NON_FINITE, "non-finite previous frame, delta or radius");
    float vp[16];
    SKYHIP_CHECK(skyhip_camera_view_projection(camera, vp, nullptr, nullptr, nullptr));
    // The local frame of VolumetricCloud::Update (:188-202), as in skyhip_cloud_assign_buffer_data.
    const V3f earth_center = {0.0f, -bottom_radius, 0.0f};
    const V3f camera_pos = load3(camera->position);
    const V3f up = normalize3(camera_pos - earth_center);
    const V3f origin = earth_center + up * bottom_radius;
    const V3f fr = normalize3(cross3(up, V3f{1.0f, 0.0f, 0.0f}));
    const V3f ri = normalize3(cross3(up, fr));
    const float model[16] = {fr.x, fr.y, fr.z, 0.0f, ri.x, ri.y, ri.z, 0.0f,
                             up.x, up.y, up.z, 0.0f, origin.x, origin.y, origin.z, 1.0f};
    float inv_model[16], mvp[16];
    if (!mat_inverse(model, inv_model)) return fail(SKYHIP_ERR_NON_FINITE, "degenerate local cloud frame");
    mat_mul(vp, model, mvp);
    if (previous_mvp) {
        // :204-214, 234: uReprojectMat = previous MVP * translate(delta_local.xy), delta_local the
        // camera's move in the local frame plus the material's additional offset.
        const V3f dw = camera_pos - load3(previous_camera_position);
        float delta[16];
        mat_identity(delta);
        delta[12] = inv_model[0] * dw.x + inv_model[4] * dw.y + inv_model[8] * dw.z;
        delta[13] = inv_model[1] * dw.x + inv_model[5] * dw.y + inv_model[9] * dw.z;
        if (additional_delta) {
            delta[12] += additional_delta[0];
            delta[13] += additional_delta[1];
        }
        mat_mul(previous_mvp, delta, common->uReprojectMat);
    } else {
        // No previous frame: this frame's MVP (a static camera).
        std::memcpy(common->uReprojectMat, mvp, sizeof(mvp));
    }
    if (out_mvp) std::memcpy(out_mvp, mvp, sizeof(mvp));
    return SKYHIP_OK;
}

extern "C" int skyhip_cloud_sample_info(float repeat_size, int32_t tex_max_width, float fovy_degrees,
                                        float viewport_w, float viewport_h, double offset_x, double offset_y,
                                        SkyhipSampleInfo* out) {
    if (out == nullptr) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL output");
    if (!(repeat_size > 0.0f) || tex_max_width <= 0 || !(viewport_w > 0.0f) || !(viewport_h > 0.0f))
        return fail(SKYHIP_ERR_INVALID_SIZE, "repeat size, texture width and viewport must be positive");
    // VolumetricCloudDefaultMaterial.cpp:88-92 and CalKLod :24-28
    out->frequency = 1.0f / repeat_size;
    double bx = offset_x / double(repeat_size), by = offset_y / double(repeat_size);
    out->bias[0] = float(bx - std::floor(bx));
    out->bias[1] = float(by - std::floor(by));
    float tan_half_fovy = std::tan(radians(fovy_degrees) * 0.5f);
    out->k_lod = float(tex_max_width) * tan_half_fovy / (repeat_size * std::fmin(viewport_w, viewport_h));
    return SKYHIP_OK;
}

// std140 layout checks (the blocks are the drop-in interface).
static_assert(sizeof(SkyhipAtmosphereBufferData) == 128, "AtmosphereBufferData is 128 bytes");
static_assert(offsetof(SkyhipAtmosphereBufferData, bottom_radius) == 112, "std140 offset");
static_assert(sizeof(SkyhipAtmosphereRenderBufferData) == 320, "AtmosphereRenderBufferData is 320 bytes");
static_assert(offsetof(SkyhipAtmosphereRenderBufferData, inv_view_projection) == 112, "std140 offset");
static_assert(offsetof(SkyhipAtmosphereRenderBufferData, padding00) == 240, "std140 offset");
static_assert(offsetof(SkyhipAtmosphereRenderBufferData, uCloudShadowMapMat) == 256, "std140 offset");
static_assert(sizeof(SkyhipCloudCommonBufferData) == 384, "VolumetricCloudCommonBufferData is 384 bytes");
static_assert(offsetof(SkyhipCloudCommonBufferData, uCameraPos) == 320, "std140 offset");
static_assert(offsetof(SkyhipCloudCommonBufferData, uSunDirection) == 352, "std140 offset");
static_assert(offsetof(SkyhipCloudCommonBufferData, uEarthRadius) == 380, "std140 offset");
static_assert(sizeof(SkyhipCloudBufferData) == 64, "VolumetricCloudBufferData is 64 bytes");
static_assert(offsetof(SkyhipCloudBufferData, uEnvColorScale) == 16, "std140 offset");
static_assert(offsetof(SkyhipCloudBufferData, uEnvSunHeightCurveExp) == 60, "std140 offset");
static_assert(sizeof(SkyhipCloudMaterialCommonBufferData) == 64, "material common block is 64 bytes");
static_assert(sizeof(SkyhipCloudMaterial1BufferData) == 32, "material1 block is 32 bytes");
static_assert(sizeof(SkyhipCloudMaterial0BufferData) == 16, "material0 block is 16 bytes");
static_assert(sizeof(SkyhipCloudMinimalBufferData) == 16, "minimal block is 16 bytes");
static_assert(sizeof(SkyhipCloudVoxelBufferData) == 32, "voxel block is 32 bytes");
static_assert(sizeof(SkyhipCloudWorleyDetailBufferData) == 32, "Worley detail block is 32 bytes");
static_assert(sizeof(SkyhipCloudPhase) == 16, "phase block is 16 bytes");
