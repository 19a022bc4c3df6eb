// Learned sky-view surrogate for gfx950: a 7 -> 64 -> 64 -> 64 -> 64 -> 3 MLP with
// smooth-step gated activations z * smoothstep(-3, 3, z) predicting RGB log-radiance of
// the sky along a view direction, with FP16 MFMA operands and FP32 accumulation, trained
// on the GPU by the kernels below and checked at run time against the plain integrator
// it replaces.
//
// Parameters (FP32 master copy, 13187 floats): for each layer l with fan-in K and
// fan-out N, W_l as N x K row-major (out-major), then b_l[N]; layers in order.
// Inputs (8 floats per sample, the 8th zero): 2 sqrt(h / H) - 1; the view zenith as the
// sky-view LUT's latitude coordinate, 2 x_lat - 1; sign(mu_sun) sqrt(|mu_sun|); the sun-view cosine
// mu_view mu_sun + sin_view sin_sun cos(dphi); log(rayleigh scale) / log 4,
// log(mie scale) / log 4, log(ozone scale) / log 4.  Output: log of the RGB radiance
// (solar illuminance units).
//
// Kernels:
//  - sky_radiance_kernel: the plain integrator per query (label generator and judge):
//    ComputeScatteredLuminance with the multiple-scattering term along the view ray
//    from altitude h, exactly the sky-view texel arithmetic for an explicit direction.
//  - surrogate_forward_kernel: inference.  One wave holds every weight in registers as
//    MFMA A operands and streams groups of 16 samples through v_mfma_f32_16x16x32_f16;
//    a layer's accumulators (rows = neurons, columns = samples) become the next layer's
//    B operand in place, the k order inside each 32-wide block being permuted to match
//    the accumulator layout (the weights are loaded with the same permutation).
//  - surrogate_train_kernel: forward and backward for 32 samples per workgroup with
//    LDS-staged FP16 operands (any transpose is an LDS stride), writing the workgroup's
//    gradient to its own slice (no atomics; the reduction order is fixed).  The deltas
//    are kept without the loss scale (2 w (y - t) and its back-propagation, well inside
//    FP16's range) and the scale multiplies the gradient sums.
//  - surrogate_reduce_kernel / surrogate_adam_kernel: gradient sum and Adam update.
//  - sky render kernels: per-pixel panorama evaluation (surrogate), a per-tile residual
//    check against the plain integrator at probe pixels, and the plain fallback for
//    tiles whose residual exceeds the bound.
#include <algorithm>
#include <mutex>
#include <vector>

#include "lut_texels.h"

namespace skyhip {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kParams = 13187;
constexpr int kOffW1 = 0, kOffB1 = 448, kOffW2 = 512, kOffB2 = 4608, kOffW3 = 4672, kOffB3 = 8768, kOffW4 = 8832,
              kOffB4 = 12928, kOffW5 = 12992, kOffB5 = 13184;

// The activation a(z) = z s(t), t = clamp(z / 6 + 1/2, 0, 1), s = t^2 (3 - 2 t): a gate that
// rises smoothly from 0 at z = -3 to 1 at z = 3 (a C1 function, like z sigmoid(z) in shape),
// built from multiplies and adds only.  Evaluated on two values at once in packed FP16, in
// this order, each operation rounding once to FP16:
//   z = FP16(z) toward zero (v_cvt_pkrtz_f16_f32), t = clamp(fma(z, FP16(1/6), 1/2), 0, 1),
//   v = fma(t, -2, 3), w = t v, s = t w, a = z s.
// The same function feeds the next layer in training and in inference.
__device__ __forceinline__ f16x2 activation2(float z0, float z1) {
    const f16x2 z = __builtin_bit_cast(f16x2, __builtin_amdgcn_cvt_pkrtz(z0, z1));
    const f16x2 sixth = {(_Float16)(1.0f / 6.0f), (_Float16)(1.0f / 6.0f)}, half = {0.5f16, 0.5f16};
    const f16x2 zero = {0.0f16, 0.0f16}, one = {1.0f16, 1.0f16}, m2 = {-2.0f16, -2.0f16}, three = {3.0f16, 3.0f16};
    f16x2 t = __builtin_elementwise_fma(z, sixth, half);
    t = __builtin_elementwise_min(__builtin_elementwise_max(t, zero), one);
    const f16x2 v = __builtin_elementwise_fma(t, m2, three);
    const f16x2 w = t * v;
    const f16x2 sg = t * w;
    return z * sg;
}
__device__ __forceinline__ _Float16 activation1(float z) { return activation2(z, 0.0f)[0]; }
// da/dz = s + z t (1 - t) (ds/dz = 6 t (1 - t) / 6), in FP32 from the FP32 pre-activation.
__device__ __forceinline__ float activation_grad(float z) {
    const float t = fminf(fmaxf(z * (1.0f / 6.0f) + 0.5f, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t) + z * t * (1.0f - t);
}

// SiLU z sigmoid(z) with the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32, each ~1 ulp),
// in FP32, rounded to BF16 for the next layer.
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// The two networks (same layers, same parameter layout, same inputs):
//  - NetF16Gate (SKYHIP_SURROGATE_F16_GATE): FP16 MFMA operands, the gated activation above
//    in packed FP16.  The default.
//  - NetBF16SiLU (SKYHIP_SURROGATE_BF16_SILU): BF16 MFMA operands, SiLU in FP32.
struct NetF16Gate {
    typedef _Float16 T;
    typedef f16x8 V8;
    __device__ static __forceinline__ f32x4 mfma(V8 a, V8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ void act_pair(float z0, float z1, T& a0, T& a1) {
        const f16x2 r = activation2(z0, z1);
        a0 = r[0];
        a1 = r[1];
    }
    __device__ static __forceinline__ T act(float z) { return activation1(z); }
    __device__ static __forceinline__ float act_grad(float z) { return activation_grad(z); }
};
struct NetBF16SiLU {
    typedef __bf16 T;
    typedef bf16x8 V8;
    __device__ static __forceinline__ f32x4 mfma(V8 a, V8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ void act_pair(float z0, float z1, T& a0, T& a1) {
        a0 = (__bf16)silu(z0);
        a1 = (__bf16)silu(z1);
    }
    __device__ static __forceinline__ T act(float z) { return (__bf16)silu(z); }
    __device__ static __forceinline__ float act_grad(float z) { return silu_grad(z); }
};

// Sky radiance along (mu_view, dphi) from altitude h for sun zenith cosine mu_sun.
__device__ inline F3 sky_radiance(const AtmDev& A, const float4* __restrict__ T, int tw, int th,
                                  const float4* __restrict__ MS, int mw, int mh, float h, float mu_v, float mu_s,
                                  float dphi, float steps) {
    const float r = A.rb + h;
    const float sin_v = sqrtf(fmaxf(1.0f - mu_v * mu_v, 0.0f)), sin_s = sqrtf(fmaxf(1.0f - mu_s * mu_s, 0.0f));
    RayStart s{r, h, mu_v, mu_s, mu_v * mu_s + sin_v * sin_s * cosf(dphi)};
    const bool hit = ray_hits_ground(A, r, h, mu_v);
    float dist = hit ? distance_to_bottom(A, r, h, mu_v) : distance_to_top(A, r, h, mu_v);
    start_from_space(A, s, dist);
    F3 L = f3(0.0f), Tr, Lf;
    if (dist > 0.0f) scattered_luminance<false>(A, T, tw, th, MS, mw, mh, s, dist, steps, 0.5f, L, Tr, Lf);
    return L;
}

struct SkyLuts {
    const SkyhipAtmosphereBufferData* atm;  // per atmosphere
    const float4* T;                        // per atmosphere, stride tw * th
    const float4* MS;                       // per atmosphere, stride mw * mh
    int tw, th, mw, mh;
    float steps;
    int count;
};

__device__ __forceinline__ int clamp_index(int a, int count) { return min(max(a, 0), count - 1); }

__device__ __forceinline__ F3 sky_radiance_q(const SkyLuts& L, const SkyhipSkyQuery& q) {
    const int a = clamp_index(q.atmosphere, L.count);
    const AtmDev A = make_atm(L.atm[a]);
    return sky_radiance(A, L.T + (long long)a * L.tw * L.th, L.tw, L.th, L.MS + (long long)a * L.mw * L.mh, L.mw,
                        L.mh, q.altitude, q.mu_view, q.mu_sun, q.delta_phi, L.steps);
}

__global__ __launch_bounds__(256) void sky_radiance_kernel(SkyLuts L, const SkyhipSkyQuery* __restrict__ q, int n,
                                                           float* __restrict__ out_rgb) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    F3 v = sky_radiance_q(L, q[i]);
    out_rgb[3 * i + 0] = v.x;
    out_rgb[3 * i + 1] = v.y;
    out_rgb[3 * i + 2] = v.z;
}

// The surrogate inputs in three parts, so that a panorama computes each part once per
// view, per row and per pixel while a single query computes all three: the functions
// below are the only definition of the inputs, which keeps the two paths bit-identical.
//
// The view zenith and azimuth enter in the sky-view LUT's coordinates
// (GetSkyViewTextureUvFromCosLatLon, AtmosphereRenderer.glsl:113-132): latitude measured
// from the horizon with square-root spacing on both sides; the relative azimuth enters as
// the sun-view cosine it implies (monotonic in delta_phi for fixed zenith angles, the
// argument of both phase functions).  The radiance is smooth in them.
struct FeatureView {      // camera and atmosphere
    float f0, f2, f4, f5, f6;  // altitude, sun zenith, log scales / ln 4
    float mu_s, horizon_up, horizon_down, sin_s;
};
struct FeatureRow {       // one view zenith of one view
    float f1, mu_v, sin_v;
};

__device__ inline FeatureView feature_view(int atmosphere, float altitude, float mu_sun,
                                           const float* __restrict__ log_scales, float thickness, int count) {
    const float inv_log4 = 0.72134752044448170368f;  // 1 / ln 4
    const int a = clamp_index(atmosphere, count);
    FeatureView v;
    v.f0 = 2.0f * sqrtf(fmaxf(altitude / thickness, 0.0f)) - 1.0f;
    const float rb = 6360.0f;  // the preset's ground radius (the horizon angle's scale)
    const float h = fmaxf(altitude, 0.0f), r = rb + h;
    v.horizon_down = acosf(sqrtf(h * (r + rb)) / r);
    v.horizon_up = kPiF - v.horizon_down;
    // The sun zenith as sign(mu_s) sqrt(|mu_s|): the square root stretches the low-sun range,
    // where the sunlight's path through the atmosphere, and so the sky's colour, changes fastest.
    v.mu_s = mu_sun;
    v.f2 = copysignf(sqrtf(fabsf(mu_sun)), mu_sun);
    v.sin_s = sqrtf(fmaxf(1.0f - mu_sun * mu_sun, 0.0f));
    v.f4 = log_scales[3 * a + 0] * inv_log4;
    v.f5 = log_scales[3 * a + 1] * inv_log4;
    v.f6 = log_scales[3 * a + 2] * inv_log4;
    return v;
}

__device__ inline FeatureRow feature_row(const FeatureView& v, float mu_view) {
    FeatureRow r;
    const float lat = acosf(fminf(fmaxf(mu_view, -1.0f), 1.0f));
    const float x_lat = lat < v.horizon_up ? 0.5f - 0.5f * sqrtf(fmaxf(1.0f - lat / v.horizon_up, 0.0f))
                                           : 0.5f + 0.5f * sqrtf(fmaxf((lat - v.horizon_up) / v.horizon_down, 0.0f));
    r.f1 = 2.0f * x_lat - 1.0f;
    r.mu_v = mu_view;
    r.sin_v = sqrtf(fmaxf(1.0f - mu_view * mu_view, 0.0f));
    return r;
}

// The 7 inputs (the 8th is 0) from the view and row parts and cos(delta_phi).
__device__ inline void features_from(const FeatureView& v, const FeatureRow& r, float cos_dphi, float f[8]) {
    f[0] = v.f0;
    f[1] = r.f1;
    f[2] = v.f2;
    f[3] = r.mu_v * v.mu_s + r.sin_v * v.sin_s * cos_dphi;
    f[4] = v.f4;
    f[5] = v.f5;
    f[6] = v.f6;
    f[7] = 0.0f;
}

// The 7 surrogate inputs of a query (the 8th is 0).
__device__ inline void features(const SkyhipSkyQuery& q, const float* __restrict__ log_scales, float thickness,
                                int count, float f[8]) {
    const FeatureView v = feature_view(q.atmosphere, q.altitude, q.mu_sun, log_scales, thickness, count);
    features_from(v, feature_row(v, q.mu_view), cosf(q.delta_phi), f);
}

__global__ __launch_bounds__(256) void features_kernel(const SkyhipSkyQuery* __restrict__ q, int n,
                                                       const float* __restrict__ log_scales, float thickness,
                                                       int count, float* __restrict__ out) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float f[8];
    features(q[i], log_scales, thickness, count, f);
    for (int k = 0; k < 8; ++k) out[8 * i + k] = f[k];
}

// ---------------------------------------------------------------------------
// Inference: weights in registers, 16 samples per MFMA column block.
// ---------------------------------------------------------------------------
// Neuron feeding k slot (8 q + i) of 32-wide block b when the B operand is built
// from accumulator tiles 2b and 2b + 1 (lane group q holds rows 4q .. 4q + 3 of each).
__device__ __forceinline__ int permuted_neuron(int b, int q, int i) {
    return 32 * b + (i < 4 ? 4 * q + i : 16 + 4 * q + (i - 4));
}

template <class Net>
struct MlpRegs {
    typename Net::V8 w1[4];
    typename Net::V8 wh[3][4][2];
    typename Net::V8 w5[2];
};

template <class Net>
__device__ inline void load_mlp(const float* __restrict__ p, MlpRegs<Net>& R, int lane) {
    typedef typename Net::T T;
    const int q = lane >> 4, r = lane & 15;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 8; ++i) {
            int k = 8 * q + i;
            R.w1[j][i] = (T)(k < 7 ? p[kOffW1 + (16 * j + r) * 7 + k] : 0.0f);
        }
    const int offs[3] = {kOffW2, kOffW3, kOffW4};
    for (int L = 0; L < 3; ++L)
        for (int j = 0; j < 4; ++j)
            for (int b = 0; b < 2; ++b)
                for (int i = 0; i < 8; ++i)
                    R.wh[L][j][b][i] = (T)p[offs[L] + (16 * j + r) * 64 + permuted_neuron(b, q, i)];
    for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 8; ++i)
            R.w5[b][i] = (T)(r < 3 ? p[kOffW5 + r * 64 + permuted_neuron(b, q, i)] : 0.0f);
}

// B operands of the next layer from the four accumulator tiles (bias included): the
// activation of accumulator items (2 i, 2 i + 1), two at a time.
template <class Net>
__device__ __forceinline__ void activate(const f32x4 c[4], typename Net::V8 h[2]) {
    typedef typename Net::T T;
    for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 2; ++i) {
            T l0, l1, h0, h1;
            Net::act_pair(c[2 * b][2 * i], c[2 * b][2 * i + 1], l0, l1);
            Net::act_pair(c[2 * b + 1][2 * i], c[2 * b + 1][2 * i + 1], h0, h1);
            h[b][2 * i] = l0;
            h[b][2 * i + 1] = l1;
            h[b][4 + 2 * i] = h0;
            h[b][4 + 2 * i + 1] = h1;
        }
}

// Forward pass of 16 samples (the lane's column); x: this lane's input operand.
// Returns the output accumulator (lanes 0..15 hold rows 0..2 in items 0..2).
template <class Net>
__device__ __forceinline__ f32x4 mlp_forward16(const MlpRegs<Net>& R, const float* s_bias, typename Net::V8 x, int q) {
    f32x4 c[4];
    typename Net::V8 h[2];
    for (int j = 0; j < 4; ++j) {
        const float* b = s_bias + 16 * j + 4 * q;
        c[j] = f32x4{b[0], b[1], b[2], b[3]};
        c[j] = Net::mfma(R.w1[j], x, c[j]);
    }
    activate<Net>(c, h);
    for (int L = 0; L < 3; ++L) {
        for (int j = 0; j < 4; ++j) {
            const float* b = s_bias + 64 * (L + 1) + 16 * j + 4 * q;
            c[j] = f32x4{b[0], b[1], b[2], b[3]};
            c[j] = Net::mfma(R.wh[L][j][0], h[0], c[j]);
            c[j] = Net::mfma(R.wh[L][j][1], h[1], c[j]);
        }
        activate<Net>(c, h);
    }
    const float* b = s_bias + 256 + 4 * q;
    f32x4 y = {b[0], b[1], b[2], b[3]};
    y = Net::mfma(R.w5[0], h[0], y);
    y = Net::mfma(R.w5[1], h[1], y);
    return y;
}

// Biases b1..b4 at 64 (l - 1), b5 at 256 (3 used, zero to 271).
__device__ inline void load_biases(const float* __restrict__ p, float* s_bias) {
    for (int i = threadIdx.x; i < 272; i += blockDim.x) {
        float v = 0.0f;
        if (i < 64) v = p[kOffB1 + i];
        else if (i < 128) v = p[kOffB2 + i - 64];
        else if (i < 192) v = p[kOffB3 + i - 128];
        else if (i < 256) v = p[kOffB4 + i - 192];
        else if (i < 259) v = p[kOffB5 + i - 256];
        s_bias[i] = v;
    }
    __syncthreads();
}

template <class Net>
__device__ __forceinline__ typename Net::V8 input_operand(const float* f, bool valid, int q) {
    typename Net::V8 x;
    for (int i = 0; i < 8; ++i) x[i] = (typename Net::T)((q == 0 && valid) ? f[i] : 0.0f);
    return x;
}

template <class Net>
__global__ __launch_bounds__(256, 2) void surrogate_forward_kernel(const float* __restrict__ params,
                                                                const float* __restrict__ feats, int n,
                                                                float* __restrict__ out) {
    __shared__ float s_bias[272];
    load_biases(params, s_bias);
    const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    MlpRegs<Net> R;
    load_mlp<Net>(params, R, lane);
    const int waves = gridDim.x * (blockDim.x >> 6);
    for (int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g * 16 < n; g += waves) {
        const int s = g * 16 + r;
        float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (q == 0 && s < n)
            for (int k = 0; k < 8; ++k) f[k] = feats[8 * s + k];
        f32x4 y = mlp_forward16<Net>(R, s_bias, input_operand<Net>(f, s < n, q), q);
        if (q == 0 && s < n) {
            out[3 * s + 0] = y[0];
            out[3 * s + 1] = y[1];
            out[3 * s + 2] = y[2];
        }
    }
}

// ---------------------------------------------------------------------------
// Training: 32 samples per workgroup, LDS-staged operands.
// ---------------------------------------------------------------------------
constexpr int kTS = 32;  // samples per workgroup

// C[M x N] = A[M x K] B[K x N] with FP16 LDS operands at element strides
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn]; 16 x 16 tiles spread over
// the workgroup's waves, K a multiple of 32.  ep(m, n, value) consumes each output once.
template <class Net, class Ep>
__device__ inline void lds_gemm(const typename Net::T* A, int sam, int sak, const typename Net::T* B, int sbk, int sbn,
                                int M, int N,
                                int K, Ep ep) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int q = lane >> 4, r = lane & 15;
    const int tn_count = N / 16;
    for (int t = wave; t < (M / 16) * tn_count; t += nw) {
        const int m0 = (t / tn_count) * 16, n0 = (t % tn_count) * 16;
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = 0; k0 < K; k0 += 32) {
            typename Net::V8 a, b;
            for (int i = 0; i < 8; ++i) {
                const int k = k0 + 8 * q + i;
                a[i] = A[(m0 + r) * sam + k * sak];
                b[i] = B[k * sbk + (n0 + r) * sbn];
            }
            c = Net::mfma(a, b, c);
        }
        for (int i = 0; i < 4; ++i) ep(m0 + 4 * q + i, n0 + r, c[i]);
    }
}

template <class T>
struct TrainLds {
    T w1[64 * 32];   // [neuron][input] (inputs 7..31 zero)
    T wh[3][64 * 64];
    T w5[32 * 64];   // [output][neuron] (outputs 3..31 zero)
    T x[32 * kTS];   // [input][sample]
    float z[4][64 * kTS]; // pre-activations [neuron][sample]
    T a[4][64 * kTS];
    float y[16 * kTS];
    T d5[32 * kTS];  // output deltas 2 w (y - t) [output][sample] (rows 3..31 zero)
    T d[2][64 * kTS];  // hidden deltas (without the loss scale)
    float bias[272];
    float loss;
};

template <class Net>
__global__ __launch_bounds__(256) void surrogate_train_kernel(const float* __restrict__ params,
                                                              const float* __restrict__ feats,
                                                              const float* __restrict__ targets,
                                                              const float* __restrict__ weights, int n,
                                                              float loss_scale, float* __restrict__ partial,
                                                              float* __restrict__ partial_loss) {
    extern __shared__ __align__(16) unsigned char smem[];
    typedef typename Net::T T;
    TrainLds<T>& S = *reinterpret_cast<TrainLds<T>*>(smem);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int s0 = blockIdx.x * kTS;
    float* G = partial + (long long)blockIdx.x * kParams;
    // Weights to FP16 (zero padding), inputs, biases.
    for (int i = tid; i < 64 * 32; i += nt) {
        int m = i / 32, k = i % 32;
        S.w1[i] = (T)(k < 7 ? params[kOffW1 + m * 7 + k] : 0.0f);
    }
    const int offs[3] = {kOffW2, kOffW3, kOffW4};
    for (int L = 0; L < 3; ++L)
        for (int i = tid; i < 64 * 64; i += nt) S.wh[L][i] = (T)params[offs[L] + i];
    for (int i = tid; i < 32 * 64; i += nt) S.w5[i] = (T)(i < 3 * 64 ? params[kOffW5 + i] : 0.0f);
    for (int i = tid; i < 32 * kTS; i += nt) {
        int k = i / kTS, s = i % kTS;
        S.x[i] = (T)((k < 8 && s0 + s < n) ? feats[8 * (s0 + s) + k] : 0.0f);
        S.d5[i] = (T)0.0f;
    }
    if (tid == 0) S.loss = 0.0f;
    load_biases(params, S.bias);  // synchronises
    // Forward.
    lds_gemm<Net>(S.w1, 32, 1, S.x, kTS, 1, 64, kTS, 32, [&](int m, int s, float c) {
        float z = c + S.bias[m];
        S.z[0][m * kTS + s] = z;
        S.a[0][m * kTS + s] = Net::act(z);
    });
    __syncthreads();
    for (int L = 0; L < 3; ++L) {
        lds_gemm<Net>(S.wh[L], 64, 1, S.a[L], kTS, 1, 64, kTS, 64, [&](int m, int s, float c) {
            float z = c + S.bias[64 * (L + 1) + m];
            S.z[L + 1][m * kTS + s] = z;
            S.a[L + 1][m * kTS + s] = Net::act(z);
        });
        __syncthreads();
    }
    lds_gemm<Net>(S.w5, 64, 1, S.a[3], kTS, 1, 16, kTS, 64,
             [&](int m, int s, float c) { S.y[m * kTS + s] = c + S.bias[256 + m]; });
    __syncthreads();
    // Loss loss_scale * sum w (y - t)^2 and the output deltas 2 w (y - t); the gradient
    // sums below multiply by loss_scale.
    for (int i = tid; i < 3 * kTS; i += nt) {
        int m = i / kTS, s = i % kTS;
        float d = 0.0f, w = 0.0f;
        if (s0 + s < n) {
            d = S.y[m * kTS + s] - targets[3 * (s0 + s) + m];
            w = weights ? weights[s0 + s] : 1.0f;
        }
        atomicAdd(&S.loss, w * d * d * loss_scale);
        S.d5[m * kTS + s] = (T)(2.0f * w * d);
    }
    __syncthreads();
    // Output layer gradient: dW5 = d5 a4^T, db5 = sum d5.
    lds_gemm<Net>(S.d5, kTS, 1, S.a[3], 1, kTS, 16, 64, kTS, [&](int m, int k, float c) {
        if (m < 3) G[kOffW5 + m * 64 + k] = c * loss_scale;
    });
    for (int m = tid; m < 3; m += nt) {
        float sum = 0.0f;
        for (int s = 0; s < kTS; ++s) sum += float(S.d5[m * kTS + s]);
        G[kOffB5 + m] = sum * loss_scale;
    }
    // delta4 = (W5^T d5) * a'(z4)
    lds_gemm<Net>(S.w5, 1, 64, S.d5, kTS, 1, 64, kTS, 32, [&](int m, int s, float c) {
        S.d[0][m * kTS + s] = (T)(c * Net::act_grad(S.z[3][m * kTS + s]));
    });
    __syncthreads();
    // Hidden layers 4, 3, 2 (weights wh[2], wh[1], wh[0]).
    const int woffs[3] = {kOffW2, kOffW3, kOffW4}, boffs[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
    int cur = 0;
    for (int L = 2; L >= 0; --L) {
        const T* dl = S.d[cur];
        // dW = delta a_prev^T ; a_prev = a[L] (input of wh[L])
        lds_gemm<Net>(dl, kTS, 1, S.a[L], 1, kTS, 64, 64, kTS,
                 [&](int m, int k, float c) { G[woffs[L] + m * 64 + k] = c * loss_scale; });
        for (int m = tid; m < 64; m += nt) {
            float sum = 0.0f;
            for (int s = 0; s < kTS; ++s) sum += float(dl[m * kTS + s]);
            G[boffs[L + 1] + m] = sum * loss_scale;
        }
        // delta_prev = (W^T delta) * a'(z[L])
        T* dn = S.d[cur ^ 1];
        lds_gemm<Net>(S.wh[L], 1, 64, dl, kTS, 1, 64, kTS, 64, [&](int m, int s, float c) {
            dn[m * kTS + s] = (T)(c * Net::act_grad(S.z[L][m * kTS + s]));
        });
        __syncthreads();
        cur ^= 1;
    }
    // First layer: dW1 = delta1 x^T, db1.
    const T* d1 = S.d[cur];
    lds_gemm<Net>(d1, kTS, 1, S.x, 1, kTS, 64, 32, kTS, [&](int m, int k, float c) {
        if (k < 7) G[kOffW1 + m * 7 + k] = c * loss_scale;
    });
    for (int m = tid; m < 64; m += nt) {
        float sum = 0.0f;
        for (int s = 0; s < kTS; ++s) sum += float(d1[m * kTS + s]);
        G[kOffB1 + m] = sum * loss_scale;
    }
    __syncthreads();
    if (tid == 0) partial_loss[blockIdx.x] = S.loss;
}

// grad[p] = sum over workgroups of partial[wg][p]; loss likewise (as parameter kParams).
// A 1024-thread block covers 64 parameters x 16 ranges of workgroups: each thread sums
// its range in order, then the 16 range sums are added in order, so the result does not
// depend on scheduling.
__global__ __launch_bounds__(1024) void surrogate_reduce_kernel(const float* __restrict__ partial,
                                                                const float* __restrict__ partial_loss, int groups,
                                                                float* __restrict__ grad, float* __restrict__ loss) {
    __shared__ float s_part[16][64];
    const int lane_p = threadIdx.x & 63, part = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane_p;
    const int chunk = (groups + 15) / 16;
    const int g0 = part * chunk, g1 = min(groups, g0 + chunk);
    float s = 0.0f;
    if (p < kParams) {
        for (int g = g0; g < g1; ++g) s += partial[(long long)g * kParams + p];
    } else if (p == kParams && loss) {
        for (int g = g0; g < g1; ++g) s += partial_loss[g];
    }
    s_part[part][lane_p] = s;
    __syncthreads();
    if (part == 0) {
        float t = 0.0f;
        for (int k = 0; k < 16; ++k) t += s_part[k][lane_p];
        if (p < kParams)
            grad[p] = t;
        else if (p == kParams && loss)
            *loss = t;
    }
}

__global__ __launch_bounds__(256) void surrogate_adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                             float* __restrict__ m, float* __restrict__ v, float lr,
                                                             float beta1, float beta2, float eps, float bc1,
                                                             float bc2) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kParams) return;
    float gi = g[p];
    float mi = beta1 * m[p] + (1.0f - beta1) * gi;
    float vi = beta2 * v[p] + (1.0f - beta2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    w[p] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---------------------------------------------------------------------------
// Panorama rendering: surrogate, residual check, plain fallback.
// ---------------------------------------------------------------------------
struct SkyRenderArgs {
    SkyLuts luts;
    const SkyhipSkyView* views;  // per view
    const float* log_scales;     // per atmosphere
    float thickness;
    int width, height, tiles_x, tiles_y;
    float4* out;                 // [view][pixel] RGB radiance (w unused)
    uint32_t* flags;             // [view][tile]: 1 if the tile fell back
    uint32_t* counters;          // [0]: tiles that fell back
    float tolerance, floor;
};

// Query of pixel (x, y) of an equirectangular panorama around the camera's zenith:
// zenith angle pi (y + 0.5) / H, azimuth from the sun 2 pi (x + 0.5) / W - pi.
__device__ __forceinline__ float panorama_mu_view(int y, int h) { return cosf(kPiF * (float(y) + 0.5f) / float(h)); }
__device__ __forceinline__ float panorama_delta_phi(int x, int w) {
    return fabsf(2.0f * kPiF * (float(x) + 0.5f) / float(w) - kPiF);
}
__device__ inline SkyhipSkyQuery panorama_query(const SkyhipSkyView& v, int x, int y, int w, int h) {
    SkyhipSkyQuery q;
    q.atmosphere = v.atmosphere;
    q.altitude = v.altitude;
    q.mu_sun = v.mu_sun;
    q.mu_view = panorama_mu_view(y, h);
    q.delta_phi = panorama_delta_phi(x, w);
    q.pad[0] = q.pad[1] = q.pad[2] = 0.0f;
    return q;
}

// Surrogate pass.  Each wave owns a contiguous range of 64-pixel runs (row-major over the
// views' pixels) and takes them in order: every lane computes its pixel's network inputs
// (the view and row parts only when its view or row changes, which for widths that are
// multiples of 64 happens for all lanes together once per row) and stages them in LDS as
// FP16; the four 16-sample column blocks then run through the network (lane group 0
// supplies the input operand of each block); the log radiance goes back through LDS so
// that every lane exponentiates and stores its own pixel.  The staging is private to the
// wave, so no workgroup barrier is needed.
template <class Net>
__global__ __launch_bounds__(256, 2) void sky_render_surrogate_kernel(const float* __restrict__ params,
                                                                   SkyRenderArgs a, int n_views) {
    __shared__ float s_bias[272];
    typedef typename Net::V8 V8;
    __shared__ V8 s_in[4][64];
    __shared__ float4 s_y[4][64];
    load_biases(params, s_bias);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    MlpRegs<Net> R;
    load_mlp<Net>(params, R, lane);
    const int w = a.width, h = a.height;
    const long long per_view = (long long)w * h;
    const long long total = per_view * n_views;
    const long long runs = (total + 63) / 64;
    const long long waves = (long long)gridDim.x * (blockDim.x >> 6);
    const long long per_wave = (runs + waves - 1) / waves;
    const long long run0 = (blockIdx.x * (long long)(blockDim.x >> 6) + wave) * per_wave;
    const long long run1 = run0 + per_wave < runs ? run0 + per_wave : runs;
    // This lane's pixel of the first run, then advanced by 64 pixels per run.
    long long idx = run0 * 64 + lane;
    int view = int(idx / per_view), x, y;
    {
        const long long pix = idx - (long long)view * per_view;
        y = int(pix / w);
        x = int(pix - (long long)y * w);
    }
    int row_view = -1, row_y = -1;  // the view and row that fv / fr were computed for
    FeatureView fv = {};
    FeatureRow fr = {};
    const V8 zero8 = {};
    for (long long run = run0; run < run1; ++run) {
        const bool valid = view < n_views;
        float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (valid) {
            if (view != row_view || y != row_y) {
                const SkyhipSkyView& v = a.views[view];
                fv = feature_view(v.atmosphere, v.altitude, v.mu_sun, a.log_scales, a.thickness, a.luts.count);
                fr = feature_row(fv, panorama_mu_view(y, h));
                row_view = view;
                row_y = y;
            }
            features_from(fv, fr, cosf(panorama_delta_phi(x, w)), f);
        }
        s_in[wave][lane] = input_operand<Net>(f, valid, 0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll 1
        for (int k = 0; k < 4; ++k) {
            const V8 xin = q == 0 ? s_in[wave][16 * k + r] : zero8;
            const f32x4 yk = mlp_forward16<Net>(R, s_bias, xin, q);
            if (q == 0) s_y[wave][16 * k + r] = make_float4(yk[0], yk[1], yk[2], 0.0f);
        }
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            const float4 o = s_y[wave][lane];
            a.out[idx] = make_float4(__expf(o.x), __expf(o.y), __expf(o.z), 0.0f);
        }
        __builtin_amdgcn_wave_barrier();
        idx += 64;
        x += 64;
        while (x >= w) {
            x -= w;
            if (++y == h) {
                y = 0;
                ++view;
            }
        }
    }
}

// Residual check: probes at the 4 quarter points of each 16 x 16 tile; a tile whose
// largest |surrogate - plain| / max(plain, floor) exceeds the tolerance is flagged.  One
// thread per probe: the 4 probes of a tile are 4 adjacent lanes, reduced with shuffles.
__global__ __launch_bounds__(256) void sky_render_check_kernel(SkyRenderArgs a, int view_count) {
    const long long tiles_per_view = (long long)a.tiles_x * a.tiles_y;
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;  // probe index
    const long long tile_global = p >> 2;
    const bool live = tile_global < tiles_per_view * view_count;
    float e = 0.0f;
    if (live) {
        const int view = int(tile_global / tiles_per_view), tile = int(tile_global % tiles_per_view);
        const int tx = tile % a.tiles_x, ty = tile / a.tiles_x, i = int(p & 3);
        const int x = min(tx * 16 + 4 + 8 * (i & 1), a.width - 1);
        const int y = min(ty * 16 + 4 + 8 * (i >> 1), a.height - 1);
        const SkyhipSkyQuery q = panorama_query(a.views[view], x, y, a.width, a.height);
        const F3 ref = sky_radiance_q(a.luts, q);
        const long long idx = (long long)view * a.width * a.height + (long long)y * a.width + x;
        const float4 s = a.out[idx];
        e = fmaxf(fmaxf(fabsf(s.x - ref.x) / fmaxf(ref.x, a.floor), fabsf(s.y - ref.y) / fmaxf(ref.y, a.floor)),
                  fabsf(s.z - ref.z) / fmaxf(ref.z, a.floor));
        if (!(e == e)) e = 3.0e38f;  // NaN counts as a failure
        a.out[idx] = make_float4(ref.x, ref.y, ref.z, 0.0f);  // probes keep the plain value
    }
    e = fmaxf(e, __shfl_xor(e, 1));
    e = fmaxf(e, __shfl_xor(e, 2));
    if (live && (p & 3) == 0) {
        const unsigned f = e > a.tolerance ? 1u : 0u;
        a.flags[tile_global] = f;
        if (f) atomicAdd(&a.counters[0], 1u);
    }
}

// Plain pass: every pixel of flagged tiles (or of all tiles when all_tiles is set).
__global__ __launch_bounds__(256) void sky_render_plain_kernel(SkyRenderArgs a, int all_tiles) {
    const int tile = blockIdx.x, view = blockIdx.y;
    if (!all_tiles && a.flags[(long long)view * a.tiles_x * a.tiles_y + tile] == 0u) return;  // workgroup-uniform
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x = tx * 16 + (threadIdx.x & 15), y = ty * 16 + (threadIdx.x >> 4);
    if (x >= a.width || y >= a.height) return;
    SkyhipSkyQuery q = panorama_query(a.views[view], x, y, a.width, a.height);
    F3 v = sky_radiance_q(a.luts, q);
    a.out[(long long)view * a.width * a.height + (long long)y * a.width + x] = make_float4(v.x, v.y, v.z, 0.0f);
}

// The reference's sky path without the surrogate: the view's sky-view LUT sampled
// bilinearly at the pixel's direction (GetSkyViewTextureUvFromCosLatLon,
// AtmosphereRenderer.glsl:113-145; the LUT's longitude is measured from the sun).
__global__ __launch_bounds__(256) void sky_render_lut_kernel(SkyRenderArgs a, const float* const* __restrict__ sv,
                                                             int sw, int sh) {
    const int tile = blockIdx.x, view = blockIdx.y;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x = tx * 16 + (threadIdx.x & 15), y = ty * 16 + (threadIdx.x >> 4);
    if (x >= a.width || y >= a.height) return;
    const SkyhipSkyQuery q = panorama_query(a.views[view], x, y, a.width, a.height);
    const AtmDev A = make_atm(a.luts.atm[clamp_index(q.atmosphere, a.luts.count)]);
    const float r = A.rb + q.altitude;
    const float horizon_down = acosf(sqrtf(q.altitude * (r + A.rb)) / r);
    const float horizon_up = kPiF - horizon_down;
    const float lat = acosf(fminf(fmaxf(q.mu_view, -1.0f), 1.0f));
    const float x_lat = lat < horizon_up ? 0.5f - 0.5f * sqrtf(fmaxf(1.0f - lat / horizon_up, 0.0f))
                                         : 0.5f + 0.5f * sqrtf(fmaxf((lat - horizon_up) / horizon_down, 0.0f));
    const float x_lon = sqrtf(fmaxf(0.5f - 0.5f * cosf(q.delta_phi), 0.0f));
    F3 L = lut_bilinear(reinterpret_cast<const float4*>(sv[view]), sw, sh, x_lon * float(sw - 1),
                        x_lat * float(sh - 1));
    a.out[(long long)view * a.width * a.height + (long long)y * a.width + x] = make_float4(L.x, L.y, L.z, 0.0f);
}

__global__ void clear_u32_kernel(uint32_t* p, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

int check_luts(const SkyhipSkyLuts* l) {
    if (!l || !l->atmospheres || !l->transmittance || !l->multiscattering)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "sky LUT set has a NULL pointer");
    if (l->transmittance_width < 2 || l->transmittance_height < 2 || l->multiscattering_width < 2 ||
        l->multiscattering_height < 2 || l->transmittance_width > 4096 || l->transmittance_height > 4096 ||
        l->multiscattering_width > 4096 || l->multiscattering_height > 4096)
        return fail(SKYHIP_ERR_INVALID_SIZE, "LUT sizes outside [2, 4096]");
    if (!(l->steps > 0.0f && l->steps <= 4096.0f)) return fail(SKYHIP_ERR_INVALID_SIZE, "steps outside (0, 4096]");
    if (l->atmosphere_count < 1) return fail(SKYHIP_ERR_INVALID_SIZE, "atmosphere_count must be >= 1");
    return SKYHIP_OK;
}

SkyLuts make_luts(const SkyhipSkyLuts* l) {
    SkyLuts s;
    s.atm = l->atmospheres;
    s.T = reinterpret_cast<const float4*>(l->transmittance);
    s.MS = reinterpret_cast<const float4*>(l->multiscattering);
    s.tw = l->transmittance_width;
    s.th = l->transmittance_height;
    s.mw = l->multiscattering_width;
    s.mh = l->multiscattering_height;
    s.steps = l->steps;
    s.count = l->atmosphere_count;
    return s;
}

}  // namespace skyhip

using namespace skyhip;

extern "C" int skyhip_surrogate_parameter_count(void) { return kParams; }

extern "C" int skyhip_sky_radiance(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyQuery* queries,
                                   int32_t count, float* out_rgb) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(check_luts(luts));
    if (!queries || !out_rgb) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count < 0) return fail(SKYHIP_ERR_INVALID_SIZE, "negative count");
    if (count == 0) return SKYHIP_OK;
    hipLaunchKernelGGL(sky_radiance_kernel, dim3(div_up(count, 256)), dim3(256), 0, ctx->stream, make_luts(luts),
                       queries, count, out_rgb);
    SKYHIP_LAUNCH_CHECK("sky_radiance_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_surrogate_features(SkyhipContext* ctx, const SkyhipSkyQuery* queries, int32_t count,
                                         const float* log_scales, int32_t atmosphere_count, float thickness,
                                         float* out_features) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!queries || !log_scales || !out_features) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count < 0) return fail(SKYHIP_ERR_INVALID_SIZE, "negative count");
    if (!(thickness > 0.0f)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "thickness must be > 0");
    if (atmosphere_count < 1) return fail(SKYHIP_ERR_INVALID_SIZE, "atmosphere_count must be >= 1");
    if (count == 0) return SKYHIP_OK;
    hipLaunchKernelGGL(features_kernel, dim3(div_up(count, 256)), dim3(256), 0, ctx->stream, queries, count,
                       log_scales, thickness, atmosphere_count, out_features);
    SKYHIP_LAUNCH_CHECK("features_kernel");
    return SKYHIP_OK;
}

static int check_network(int32_t network) {
    if (network != SKYHIP_SURROGATE_F16_GATE && network != SKYHIP_SURROGATE_BF16_SILU)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "unknown surrogate network kind");
    return SKYHIP_OK;
}

extern "C" int skyhip_surrogate_forward(SkyhipContext* ctx, int32_t network, const float* params,
                                        const float* features, int32_t count, float* out_log_radiance) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(check_network(network));
    if (!params || !features || !out_log_radiance) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count < 0) return fail(SKYHIP_ERR_INVALID_SIZE, "negative count");
    if (count == 0) return SKYHIP_OK;
    const int groups = int(div_up(count, 16));
    const int blocks = std::max(1, std::min(int(div_up(groups, 4)), ctx->cu_count * 8));
    if (network == SKYHIP_SURROGATE_F16_GATE)
        hipLaunchKernelGGL(surrogate_forward_kernel<NetF16Gate>, dim3(blocks), dim3(256), 0, ctx->stream, params,
                           features, count, out_log_radiance);
    else
        hipLaunchKernelGGL(surrogate_forward_kernel<NetBF16SiLU>, dim3(blocks), dim3(256), 0, ctx->stream, params,
                           features, count, out_log_radiance);
    SKYHIP_LAUNCH_CHECK("surrogate_forward_kernel");
    return SKYHIP_OK;
}

extern "C" size_t skyhip_surrogate_workspace_bytes(int32_t count) {
    if (count <= 0) return 0;
    size_t groups = size_t((count + kTS - 1) / kTS);
    return groups * kParams * sizeof(float) + groups * sizeof(float);
}

namespace skyhip {

// Launches the training kernels for `count` samples with the per-workgroup partial gradients
// (div_up(count, kTS) x kParams floats, then one loss per workgroup) at `partial`; every
// workgroup writes its whole slice, so the buffer needs no clearing.
static int gradient_launch(SkyhipContext* ctx, int32_t network, const float* params, const float* features,
                           const float* targets, const float* weights, int32_t count, float loss_scale,
                           float* out_grad, float* out_loss, float* partial) {
    // The training kernels' LDS footprint exceeds the 64 KiB default (both element types are
    // two bytes, so both layouts have the same size).
    static_assert(sizeof(TrainLds<_Float16>) == sizeof(TrainLds<__bf16>), "LDS layouts differ");
    const int lds_bytes = int(sizeof(TrainLds<_Float16>));
    static bool lds_ready = false;
    if (!lds_ready) {
        SKYHIP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&surrogate_train_kernel<NetF16Gate>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes));
        SKYHIP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&surrogate_train_kernel<NetBF16SiLU>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes));
        lds_ready = true;
    }
    const int groups = div_up(count, kTS);
    float* partial_loss = partial + (size_t)groups * kParams;
    if (network == SKYHIP_SURROGATE_F16_GATE)
        hipLaunchKernelGGL(surrogate_train_kernel<NetF16Gate>, dim3(groups), dim3(256), lds_bytes, ctx->stream,
                           params, features, targets, weights, count, loss_scale, partial, partial_loss);
    else
        hipLaunchKernelGGL(surrogate_train_kernel<NetBF16SiLU>, dim3(groups), dim3(256), lds_bytes, ctx->stream,
                           params, features, targets, weights, count, loss_scale, partial, partial_loss);
    SKYHIP_LAUNCH_CHECK("surrogate_train_kernel");
    hipLaunchKernelGGL(surrogate_reduce_kernel, dim3(div_up(kParams + 1, 64)), dim3(1024), 0, ctx->stream, partial,
                       partial_loss, groups, out_grad, out_loss);
    SKYHIP_LAUNCH_CHECK("surrogate_reduce_kernel");
    return SKYHIP_OK;
}

static int check_gradient_args(SkyhipContext* ctx, int32_t network, const float* params, const float* features,
                               const float* targets, const float* out_grad, int32_t count, float loss_scale) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(check_network(network));
    if (!params || !features || !targets || !out_grad) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count < 1 || count > (1 << 26)) return fail(SKYHIP_ERR_INVALID_SIZE, "count outside [1, 2^26]");
    if (!std::isfinite(loss_scale)) return fail(SKYHIP_ERR_NON_FINITE, "non-finite loss scale");
    return SKYHIP_OK;
}

// The allocation-free gradient's partial gradients: one copy of this array per GPU, held until the
// process ends (13.5 MB: SKYHIP_SURROGATE_NOALLOC_MAX_COUNT / 32 workgroups x (13,187 + 1) floats).
// Host code never uses the array's name as an address: surrogate_scratch_setup resolves each
// device's copy with hipGetSymbolAddress when a context is created on it.
constexpr int kNoAllocMaxGroups = SKYHIP_SURROGATE_NOALLOC_MAX_COUNT / kTS;
static_assert(SKYHIP_SURROGATE_NOALLOC_MAX_COUNT % kTS == 0, "whole workgroups");
static __device__ __align__(256) float scratch[size_t(kNoAllocMaxGroups) * (kParams + 1)];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static float* g_scratch[kMaxScratchDevices] = {};

// Called by skyhip_create with `device` current: resolves that device's copy of `scratch` once.
int surrogate_scratch_setup(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "device %d outside [0, %d)", device, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return SKYHIP_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    if (ptr == nullptr) return fail(SKYHIP_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave NULL");
    g_scratch[device] = static_cast<float*>(ptr);
    return SKYHIP_OK;
}

static float* scratch_for(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch[device];
}

}  // namespace skyhip

extern "C" int skyhip_surrogate_gradient(SkyhipContext* ctx, int32_t network, const float* params,
                                         const float* features, const float* targets, const float* weights,
                                         int32_t count, float loss_scale, float* out_grad, float* out_loss,
                                         void* workspace, size_t workspace_bytes) {
    SKYHIP_CHECK(check_gradient_args(ctx, network, params, features, targets, out_grad, count, loss_scale));
    if (!workspace) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (workspace_bytes < skyhip_surrogate_workspace_bytes(count))
        return fail(SKYHIP_ERR_INVALID_SIZE, "workspace smaller than skyhip_surrogate_workspace_bytes(count)");
    return gradient_launch(ctx, network, params, features, targets, weights, count, loss_scale, out_grad, out_loss,
                           reinterpret_cast<float*>(workspace));
}

extern "C" int skyhip_surrogate_gradient_noalloc(SkyhipContext* ctx, int32_t network, const float* params,
                                                 const float* features, const float* targets, const float* weights,
                                                 int32_t count, float loss_scale, float* out_grad,
                                                 float* out_loss) {
    SKYHIP_CHECK(check_gradient_args(ctx, network, params, features, targets, out_grad, count, loss_scale));
    if (count > SKYHIP_SURROGATE_NOALLOC_MAX_COUNT)
        return fail(SKYHIP_ERR_INVALID_SIZE, "count %d above the allocation-free maximum %d", count,
                    SKYHIP_SURROGATE_NOALLOC_MAX_COUNT);
    float* partial = scratch_for(ctx->device);
    if (partial == nullptr)
        return fail(SKYHIP_ERR_DEVICE, "the allocation-free workspace of device %d was not resolved", ctx->device);
    return gradient_launch(ctx, network, params, features, targets, weights, count, loss_scale, out_grad, out_loss,
                           partial);
}

extern "C" int skyhip_surrogate_adam(SkyhipContext* ctx, float* params, const float* grad, float* m, float* v,
                                     float lr, float beta1, float beta2, float eps, int32_t step) {
    SKYHIP_CHECK(check_ctx(ctx));
    if (!params || !grad || !m || !v) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (step < 1) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "Adam steps count from 1");
    if (!(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "bad Adam hyper-parameters");
    const float bc1 = 1.0f - powf(beta1, float(step)), bc2 = 1.0f - powf(beta2, float(step));
    hipLaunchKernelGGL(surrogate_adam_kernel, dim3(div_up(kParams, 256)), dim3(256), 0, ctx->stream, params, grad,
                       m, v, lr, beta1, beta2, eps, bc1, bc2);
    SKYHIP_LAUNCH_CHECK("surrogate_adam_kernel");
    return SKYHIP_OK;
}

static int render_args(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyView* views, int32_t n_views,
                       const float* log_scales, float thickness, int32_t width, int32_t height, float* out_rgba,
                       uint32_t* tile_flags, uint32_t* counters, float tolerance, float floor_, SkyRenderArgs& a) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(check_luts(luts));
    if (!views || !out_rgba || !tile_flags || !counters) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n_views < 1 || n_views > 65535) return fail(SKYHIP_ERR_INVALID_SIZE, "view count outside [1, 65535]");
    if (width < 1 || height < 1 || width > 16384 || height > 16384)
        return fail(SKYHIP_ERR_INVALID_SIZE, "image size outside [1, 16384]");
    if (!(thickness > 0.0f)) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "thickness must be > 0");
    a.luts = make_luts(luts);
    a.views = views;
    a.log_scales = log_scales;
    a.thickness = thickness;
    a.width = width;
    a.height = height;
    a.tiles_x = div_up(width, 16);
    a.tiles_y = div_up(height, 16);
    a.out = reinterpret_cast<float4*>(out_rgba);
    a.flags = tile_flags;
    a.counters = counters;
    a.tolerance = tolerance;
    a.floor = floor_;
    return SKYHIP_OK;
}

extern "C" int skyhip_sky_render_surrogate(SkyhipContext* ctx, int32_t network, const float* params,
                                           const SkyhipSkyLuts* luts, const SkyhipSkyView* views, int32_t view_count,
                                           const float* log_scales, float thickness, int32_t width, int32_t height,
                                           float tolerance, float radiance_floor, float* out_rgba,
                                           unsigned* tile_flags, unsigned* counters) {
    SKYHIP_CHECK(check_ctx(ctx));
    SKYHIP_CHECK(check_network(network));
    SkyRenderArgs a;
    SKYHIP_CHECK(render_args(ctx, luts, views, view_count, log_scales, thickness, width, height, out_rgba, tile_flags,
                             counters, tolerance, radiance_floor, a));
    if (!params || !log_scales) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(tolerance > 0.0f) || !(radiance_floor > 0.0f))
        return fail(SKYHIP_ERR_INVALID_ARGUMENT, "tolerance and radiance floor must be > 0");
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(clear_u32_kernel, dim3(1), dim3(256), 0, st, counters, 1);
    // One wave per 64-pixel run, grid-strided; two workgroups per CU is what the kernel's
    // registers (the weights) leave room for, so each wave loads the weights once.
    const long long runs = ((long long)width * height * view_count + 63) / 64;
    const int blocks = int(std::max(1LL, std::min((runs + 3) / 4, (long long)ctx->cu_count * 2)));
    if (network == SKYHIP_SURROGATE_F16_GATE)
        hipLaunchKernelGGL(sky_render_surrogate_kernel<NetF16Gate>, dim3(blocks), dim3(256), 0, st, params, a,
                           view_count);
    else
        hipLaunchKernelGGL(sky_render_surrogate_kernel<NetBF16SiLU>, dim3(blocks), dim3(256), 0, st, params, a,
                           view_count);
    dim3 tiles(a.tiles_x * a.tiles_y, view_count);
    const long long probes = 4LL * a.tiles_x * a.tiles_y * view_count;
    hipLaunchKernelGGL(sky_render_check_kernel, dim3(unsigned((probes + 255) / 256)), dim3(256), 0, st, a, view_count);
    hipLaunchKernelGGL(sky_render_plain_kernel, tiles, dim3(256), 0, st, a, 0);
    SKYHIP_LAUNCH_CHECK("sky render (surrogate)");
    return SKYHIP_OK;
}

extern "C" int skyhip_sky_render_plain(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyView* views,
                                       int32_t view_count, int32_t width, int32_t height, float* out_rgba) {
    SkyRenderArgs a;
    uint32_t dummy_flags = 0, dummy_counter = 0;  // never dereferenced with all_tiles set
    SKYHIP_CHECK(render_args(ctx, luts, views, view_count, nullptr, 1.0f, width, height, out_rgba, &dummy_flags,
                             &dummy_counter, 1.0f, 1.0f, a));
    a.flags = nullptr;
    a.counters = nullptr;
    dim3 tiles(a.tiles_x * a.tiles_y, view_count);
    hipLaunchKernelGGL(sky_render_plain_kernel, tiles, dim3(256), 0, ctx->stream, a, 1);
    SKYHIP_LAUNCH_CHECK("sky_render_plain_kernel");
    return SKYHIP_OK;
}

extern "C" int skyhip_sky_render_lut(SkyhipContext* ctx, const SkyhipSkyLuts* luts, const SkyhipSkyView* views,
                                     int32_t view_count, const float* const* sky_view_luminance,
                                     int32_t sky_view_width, int32_t sky_view_height, int32_t width, int32_t height,
                                     float* out_rgba) {
    SkyRenderArgs a;
    uint32_t dummy_flags = 0, dummy_counter = 0;
    SKYHIP_CHECK(render_args(ctx, luts, views, view_count, nullptr, 1.0f, width, height, out_rgba, &dummy_flags,
                             &dummy_counter, 1.0f, 1.0f, a));
    if (!sky_view_luminance) return fail(SKYHIP_ERR_INVALID_ARGUMENT, "NULL sky-view LUT array");
    if (sky_view_width < 2 || sky_view_height < 2 || sky_view_width > 8192 || sky_view_height > 8192)
        return fail(SKYHIP_ERR_INVALID_SIZE, "sky-view size outside [2, 8192]");
    a.flags = nullptr;
    a.counters = nullptr;
    dim3 tiles(a.tiles_x * a.tiles_y, view_count);
    hipLaunchKernelGGL(sky_render_lut_kernel, tiles, dim3(256), 0, ctx->stream, a, sky_view_luminance, sky_view_width,
                       sky_view_height);
    SKYHIP_LAUNCH_CHECK("sky_render_lut_kernel");
    return SKYHIP_OK;
}
