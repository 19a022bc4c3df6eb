// Shared host/device helpers for the skyhip library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <string>

#include "../../include/skyhip.h"

namespace skyhip {

// ---------------------------------------------------------------------------
// Error reporting across the C ABI: a per-thread message, never an exception.
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int check_hip(hipError_t e, const char* what);

#define SKYHIP_CHECK(call)                                 \
    do {                                                   \
        int _rc = (call);                                  \
        if (_rc != SKYHIP_OK) return _rc;                  \
    } while (0)

#define SKYHIP_HIP(call)                                   \
    do {                                                   \
        hipError_t _e = (call);                            \
        if (_e != hipSuccess) return ::skyhip::check_hip(_e, #call); \
    } while (0)

// Launch-then-check: reports a failed launch with the kernel's name.
#define SKYHIP_LAUNCH_CHECK(name)                                   \
    do {                                                            \
        hipError_t _e = hipGetLastError();                          \
        if (_e != hipSuccess) return ::skyhip::check_hip(_e, name); \
    } while (0)

bool finite_block(const float* p, size_t n);
int require(bool cond, int code, const char* what);

}  // namespace skyhip

struct SkyhipContext {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux_stream = nullptr;  // fork/join partner for independent passes
    hipEvent_t fork_event = nullptr, join_event = nullptr;
    hipEvent_t events[64] = {};
    int cu_count = 0;
    int wave_size = 64;
    int clock_khz = 0;
    char name[256] = {};
    // The bound ray-scatter-visibility froxel (R16 UNORM [d][h][w]), or null: visibility 1.
    const uint16_t* froxel = nullptr;
    int froxel_w = 0, froxel_h = 0, froxel_d = 0;
    // The raymarch's dual-lobe phase (the reference's constant lobes by default).
    SkyhipCloudPhase phase = {-0.15f, 2.16f, 0.85f, 0.0f};
    // The persistent path tracer's tile counter and the brick walks' inverse majorants (device
    // memory, allocated on first use; the latter grows to the largest brick grid traced).
    unsigned* pt_queue = nullptr;
    float* pt_inv_majorants = nullptr;
    size_t pt_inv_majorants_count = 0;
};

namespace skyhip {

int check_ctx(SkyhipContext* ctx);
// Resolves `device`'s copy of the allocation-free surrogate gradient's workspace (surrogate.hip).
int surrogate_scratch_setup(int device);

// ---------------------------------------------------------------------------
// Device vector helpers (float3 arithmetic; HIP's float3 is a plain struct).
// ---------------------------------------------------------------------------
struct F3 {
    float x, y, z;
};
__host__ __device__ inline F3 f3(float a) { return {a, a, a}; }
__host__ __device__ inline F3 f3(float a, float b, float c) { return {a, b, c}; }
__host__ __device__ inline F3 f3(const float* p) { return {p[0], p[1], p[2]}; }
__host__ __device__ inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline F3 operator*(float s, F3 a) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline F3& operator+=(F3& a, F3 b) { a = a + b; return a; }
__host__ __device__ inline F3& operator*=(F3& a, F3 b) { a = a * b; return a; }
__host__ __device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline F3 cross(F3 a, F3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

__device__ inline float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ inline float fast_exp(float x) { return __expf(x); }
__device__ inline F3 fast_exp3(F3 v) { return {__expf(v.x), __expf(v.y), __expf(v.z)}; }
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline float smoothstepf(float e0, float e1, float x) {
    float t = clamp01((x - e0) / (e1 - e0));
    return t * t * (3.0f - 2.0f * t);
}

// Launch geometry helper.
inline unsigned div_up(long long a, long long b) { return unsigned((a + b - 1) / b); }

}  // namespace skyhip
