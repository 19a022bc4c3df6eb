// Test infrastructure, not product code: the CPU oracle for the SkyRendering
// hot path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it, as the checker.  The product path (skyrendering_amd/) never
// links or calls it.
//
// Scalar-templated restatement of the reference GLSL.  The FP64 instantiation
// is the parity judge; the FP32 one (OpenMP) is the timed CPU baseline.
//
// This header holds the GLSL built-in semantics the restatement relies on
// (GLSL 4.60 spec, section 8) and the GL sampler rules (GL 4.6 spec,
// section 8.14): bilinear filtering with exact weights, CLAMP_TO_EDGE /
// REPEAT / CLAMP_TO_BORDER(0) wrapping, and NEAREST_MIPMAP_NEAREST level
// selection with the c = 0.5 magnification switch.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

constexpr double kPi = 3.1415926535897932384626433832795;  // shaders/Base/Common.glsl:4

template <class S>
struct V2 {
    S x, y;
};

template <class S>
struct V3 {
    S x, y, z;
    S& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    S operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

template <class S> inline V3<S> v3(S a) { return {a, a, a}; }
template <class S> inline V3<S> v3(S a, S b, S c) { return {a, b, c}; }
template <class S> inline V3<S> operator+(V3<S> a, V3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> inline V3<S> operator-(V3<S> a, V3<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> inline V3<S> operator-(V3<S> a) { return {-a.x, -a.y, -a.z}; }
template <class S> inline V3<S> operator*(V3<S> a, V3<S> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class S> inline V3<S> operator/(V3<S> a, V3<S> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class S> inline V3<S> operator*(V3<S> a, S s) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> inline V3<S> operator*(S s, V3<S> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class S> inline V3<S> operator/(V3<S> a, S s) { return {a.x / s, a.y / s, a.z / s}; }
template <class S> inline V3<S>& operator+=(V3<S>& a, V3<S> b) { a = a + b; return a; }
template <class S> inline V3<S>& operator*=(V3<S>& a, V3<S> b) { a = a * b; return a; }
template <class S> inline V3<S>& operator*=(V3<S>& a, S s) { a = a * s; return a; }
template <class S> inline S dot(V3<S> a, V3<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> inline V3<S> cross(V3<S> a, V3<S> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class S> inline S length(V3<S> a) { return std::sqrt(dot(a, a)); }
template <class S> inline S distance(V3<S> a, V3<S> b) { return length(a - b); }
template <class S> inline V3<S> normalize(V3<S> a) { return a * (S(1) / length(a)); }
template <class S> inline V3<S> vexp(V3<S> a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }

template <class S, class T> inline V3<S> cvt3(const T* p) { return {S(p[0]), S(p[1]), S(p[2])}; }

// GLSL clamp(x, lo, hi) = min(max(x, lo), hi); NaN follows std::fmin/fmax
// (the non-NaN operand wins), which is what GPU min/max instructions do.
template <class S> inline S gclamp(S x, S lo, S hi) { return std::fmin(std::fmax(x, lo), hi); }
template <class S> inline S gmix(S x, S y, S a) { return x * (S(1) - a) + y * a; }
template <class S> inline S fract(S x) { return x - std::floor(x); }
template <class S> inline S smoothstep(S e0, S e1, S x) {
    S t = gclamp((x - e0) / (e1 - e0), S(0), S(1));
    return t * t * (S(3) - S(2) * t);
}

// GLSL mat4 * vec4 with a column-major float[16] (glm memory order).
template <class S> inline void mat4_mul(const float* m, const S v[4], S out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = S(m[0 * 4 + r]) * v[0] + S(m[1 * 4 + r]) * v[1] + S(m[2 * 4 + r]) * v[2] + S(m[3 * 4 + r]) * v[3];
}

// ProjectiveMul, shaders/Base/Common.glsl:7-10.
template <class S> inline V3<S> projective_mul(const float* m, V3<S> v) {
    S in[4] = {v.x, v.y, v.z, S(1)}, o[4];
    mat4_mul(m, in, o);
    return {o[0] / o[3], o[1] / o[3], o[2] / o[3]};
}

// ---------------------------------------------------------------------------
// Sampler emulation (GL 4.6, section 8.14).
// ---------------------------------------------------------------------------

// CLAMP_TO_EDGE / LINEAR filtering of an RGBA32F 2D texture at normalised uv.
template <class S> inline V3<S> tex2d_linear_clamp(const float* rgba, int w, int h, S u, S v) {
    S x = u * S(w) - S(0.5), y = v * S(h) - S(0.5);
    S fx = std::floor(x), fy = std::floor(y);
    S ax = x - fx, ay = y - fy;
    long ix = long(fx), iy = long(fy);
    auto cx = [w](long i) { return int(i < 0 ? 0 : (i >= w ? w - 1 : i)); };
    auto cy = [h](long i) { return int(i < 0 ? 0 : (i >= h ? h - 1 : i)); };
    int x0 = cx(ix), x1 = cx(ix + 1), y0 = cy(iy), y1 = cy(iy + 1);
    V3<S> r;
    for (int c = 0; c < 3; ++c) {
        S t00 = rgba[(size_t(y0) * w + x0) * 4 + c], t10 = rgba[(size_t(y0) * w + x1) * 4 + c];
        S t01 = rgba[(size_t(y1) * w + x0) * 4 + c], t11 = rgba[(size_t(y1) * w + x1) * 4 + c];
        r[c] = (S(1) - ax) * (S(1) - ay) * t00 + ax * (S(1) - ay) * t10 + (S(1) - ax) * ay * t01 + ax * ay * t11;
    }
    return r;
}

// CLAMP_TO_EDGE / LINEAR filtering of an RGBA32F 3D texture at normalised uvw.
template <class S> inline V3<S> tex3d_linear_clamp(const float* rgba, int w, int h, int d, S u, S v, S q) {
    S c[3] = {u * S(w) - S(0.5), v * S(h) - S(0.5), q * S(d) - S(0.5)};
    int n[3] = {w, h, d};
    int i0[3], i1[3];
    S a[3];
    for (int k = 0; k < 3; ++k) {
        S f = std::floor(c[k]);
        a[k] = c[k] - f;
        long i = long(f);
        i0[k] = int(i < 0 ? 0 : (i >= n[k] ? n[k] - 1 : i));
        i1[k] = int(i + 1 < 0 ? 0 : (i + 1 >= n[k] ? n[k] - 1 : i + 1));
    }
    V3<S> r = v3(S(0));
    for (int corner = 0; corner < 8; ++corner) {
        int xi = (corner & 1) ? i1[0] : i0[0], yi = (corner & 2) ? i1[1] : i0[1], zi = (corner & 4) ? i1[2] : i0[2];
        S wgt = ((corner & 1) ? a[0] : S(1) - a[0]) * ((corner & 2) ? a[1] : S(1) - a[1]) *
                ((corner & 4) ? a[2] : S(1) - a[2]);
        const float* t = rgba + ((size_t(zi) * h + yi) * w + xi) * 4;
        r.x += wgt * S(t[0]);
        r.y += wgt * S(t[1]);
        r.z += wgt * S(t[2]);
    }
    return r;
}

// Mip chain of a UNORM8 texture with `channels` channels, levels packed
// back to back (level n has max(1, size >> n) texels per axis).
inline size_t mip_offset(int size, int dims, int channels, int level) {
    size_t off = 0;
    for (int l = 0; l < level; ++l) {
        size_t s = size_t(std::max(1, size >> l));
        off += (dims == 3 ? s * s * s : s * s) * channels;
    }
    return off;
}

// GL NEAREST_MIPMAP_NEAREST level choice for textureLod(λ) with LINEAR
// magnification: c = 0.5.  Returns -1 for magnification (λ <= c: LINEAR on the
// base level), otherwise the level d = ceil(λ + 0.5) - 1 clamped to q.
template <class S> inline int nearest_mip_level(S lambda, int levels) {
    if (!(lambda > S(0.5))) return -1;
    S dd = std::ceil(lambda + S(0.5)) - S(1);
    int q = levels - 1;
    if (!(dd < S(q))) return q;
    return int(dd);
}

// The checker's record of how close the sampler's discrete choices came to flipping: the
// smallest |x - round(x)| / max(|x|, 1) over the texel-space coordinates x = u * size of NEAREST
// fetches (a texel choice flips where x crosses an integer; FP32 rounding of the sample position
// moves x by a few ulps of x, hence the relative measure), and the smallest distance of lambda from the
// magnification switch (lambda = 0.5) and from the boundary between two levels that the clamp
// to the last level does not hide (lambda + 0.5 an integer).  Recorded only while a record is
// active on the calling thread (ActiveMargins), so other callers pay nothing.
struct DecisionMargins {
    double texel = 1e30, lod = 1e30;
};
inline DecisionMargins*& ActiveMargins() {
    static thread_local DecisionMargins* active = nullptr;
    return active;
}
template <class S> inline void note_lod(S lambda, int levels) {
    DecisionMargins* m = ActiveMargins();
    const double l = double(lambda);
    if (!m || !std::isfinite(l)) return;
    m->lod = std::min(m->lod, std::fabs(l - 0.5));
    const double x = l + 0.5, n = std::round(x);
    if (l > 0.0 && n - 1.0 < double(levels - 1)) m->lod = std::min(m->lod, std::fabs(x - n));
}
template <class S> inline void note_texel(S x) {
    DecisionMargins* m = ActiveMargins();
    if (m) {
        const double v = double(x);
        m->texel = std::min(m->texel, std::fabs(v - std::round(v)) / std::max(std::fabs(v), 1.0));
    }
}

inline int wrap_repeat(long i, int n) {
    long m = i % n;
    return int(m < 0 ? m + n : m);
}

// textureLod on a REPEAT-wrapped, mipmapped UNORM8 2D texture with LINEAR
// magnification and NEAREST_MIPMAP_NEAREST minification. Returns the first
// `nch` channels of `channels` as UNORM values (v / 255).
template <class S>
inline void texlod2d_u8_repeat(const uint8_t* tex, int size, int levels, int channels, S u, S v, S lambda,
                               int nch, S* out) {
    note_lod(lambda, levels);
    int lvl = nearest_mip_level(lambda, levels);
    if (lvl < 0) {
        S x = u * S(size) - S(0.5), y = v * S(size) - S(0.5);
        S fx = std::floor(x), fy = std::floor(y);
        S ax = x - fx, ay = y - fy;
        int x0 = wrap_repeat(long(fx), size), x1 = wrap_repeat(long(fx) + 1, size);
        int y0 = wrap_repeat(long(fy), size), y1 = wrap_repeat(long(fy) + 1, size);
        for (int c = 0; c < nch; ++c) {
            S t00 = tex[(size_t(y0) * size + x0) * channels + c], t10 = tex[(size_t(y0) * size + x1) * channels + c];
            S t01 = tex[(size_t(y1) * size + x0) * channels + c], t11 = tex[(size_t(y1) * size + x1) * channels + c];
            out[c] = ((S(1) - ax) * (S(1) - ay) * t00 + ax * (S(1) - ay) * t10 + (S(1) - ax) * ay * t01 +
                      ax * ay * t11) / S(255);
        }
        return;
    }
    int s = std::max(1, size >> lvl);
    const uint8_t* base = tex + mip_offset(size, 2, channels, lvl);
    note_texel(u * S(s));
    note_texel(v * S(s));
    int xi = wrap_repeat(long(std::floor(u * S(s))), s), yi = wrap_repeat(long(std::floor(v * S(s))), s);
    for (int c = 0; c < nch; ++c) out[c] = S(base[(size_t(yi) * s + xi) * channels + c]) / S(255);
}

// textureLod on a mipmapped single-channel UNORM8 3D texture (cube of `size`),
// LINEAR magnification, NEAREST_MIPMAP_NEAREST minification.  `border` selects
// CLAMP_TO_BORDER with border colour 0 instead of REPEAT.
template <class S>
inline S texlod3d_u8(const uint8_t* tex, const int dim[3], int levels, S u, S v, S w, S lambda, bool border) {
    note_lod(lambda, levels);
    int lvl = nearest_mip_level(lambda, levels);
    S coord[3] = {u, v, w};
    if (lvl < 0) {
        int i0[3], i1[3];
        bool in0[3], in1[3];
        S a[3];
        for (int k = 0; k < 3; ++k) {
            S c = coord[k] * S(dim[k]) - S(0.5);
            S f = std::floor(c);
            a[k] = c - f;
            long i = long(f);
            if (border) {
                in0[k] = i >= 0 && i < dim[k];
                in1[k] = i + 1 >= 0 && i + 1 < dim[k];
                i0[k] = int(i);
                i1[k] = int(i + 1);
            } else {
                in0[k] = in1[k] = true;
                i0[k] = wrap_repeat(i, dim[k]);
                i1[k] = wrap_repeat(i + 1, dim[k]);
            }
        }
        S r = 0;
        for (int corner = 0; corner < 8; ++corner) {
            bool bx = corner & 1, by = corner & 2, bz = corner & 4;
            S wgt = (bx ? a[0] : S(1) - a[0]) * (by ? a[1] : S(1) - a[1]) * (bz ? a[2] : S(1) - a[2]);
            bool inside = (bx ? in1[0] : in0[0]) && (by ? in1[1] : in0[1]) && (bz ? in1[2] : in0[2]);
            S t = 0;
            if (inside) {
                int xi = bx ? i1[0] : i0[0], yi = by ? i1[1] : i0[1], zi = bz ? i1[2] : i0[2];
                t = S(tex[(size_t(zi) * dim[1] + yi) * dim[0] + xi]);
            }
            r += wgt * t;
        }
        return r / S(255);
    }
    int d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::max(1, dim[k] >> lvl);
    size_t off = 0;
    for (int l = 0; l < lvl; ++l) {
        size_t a = std::max(1, dim[0] >> l), b = std::max(1, dim[1] >> l), c = std::max(1, dim[2] >> l);
        off += a * b * c;
    }
    int idx[3];
    for (int k = 0; k < 3; ++k) {
        note_texel(coord[k] * S(d[k]));
        long i = long(std::floor(coord[k] * S(d[k])));
        if (border) {
            if (i < 0 || i >= d[k]) return S(0);
            idx[k] = int(i);
        } else {
            idx[k] = wrap_repeat(i, d[k]);
        }
    }
    return S(tex[off + (size_t(idx[2]) * d[1] + idx[1]) * d[0] + idx[0]]) / S(255);
}

}  // namespace oracle
