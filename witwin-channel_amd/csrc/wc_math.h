// Device math for the MI355X propagation kernels: 3-vectors, complex numbers and
// forward-mode dual numbers, all templated on the scalar so one path evaluator
// serves the float forward pass, the double checks and the tangent (VJP) pass.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#define WC_HD __host__ __device__ __forceinline__

namespace wc {

constexpr double kC0 = 299792458.0;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
// Numerical guards of the field-transport formulas: EPS bounds a length or
// norm away from zero, SMALL_EPS decides when a vector is too short to
// normalize and clamps material/angle inputs (reference tests/ad/_reference_fields.py:13-14).
constexpr float kEpsF = 1.0e-10f;
constexpr float kSmallEpsF = 1.0e-6f;

// ---------------------------------------------------------------- dual numbers
template <int N>
struct Dual {
    float v;
    float d[N];
};

template <typename R> struct ScalarTraits;
template <> struct ScalarTraits<float> { static constexpr int N = 0; };
template <> struct ScalarTraits<double> { static constexpr int N = 0; };
template <int N> struct ScalarTraits<Dual<N>> { static constexpr int N_ = N; static constexpr int N2 = N; };

WC_HD float val(float x) { return x; }
WC_HD double val(double x) { return x; }
template <int N> WC_HD float val(const Dual<N>& x) { return x.v; }

template <typename R> WC_HD R cst(float x);
template <> WC_HD float cst<float>(float x) { return x; }
template <> WC_HD double cst<double>(float x) { return (double)x; }

template <int N>
WC_HD Dual<N> make_dual(float v) {
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = 0.0f;
    return r;
}
template <int N>
WC_HD Dual<N> seed(float v, int which) {
    Dual<N> r = make_dual<N>(v);
    if (which >= 0 && which < N) r.d[which] = 1.0f;
    return r;
}

// lift a plain value into R
template <typename R> struct Lift { WC_HD static R from(float x) { return (R)x; } };
template <int N> struct Lift<Dual<N>> { WC_HD static Dual<N> from(float x) { return make_dual<N>(x); } };
template <typename R> WC_HD R lift(float x) { return Lift<R>::from(x); }

template <int N> WC_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N> WC_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N> WC_HD Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r; r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N> WC_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N> WC_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; const float inv = 1.0f / b.v; r.v = a.v * inv;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
template <int N> WC_HD Dual<N> operator+(const Dual<N>& a, float b) { Dual<N> r = a; r.v += b; return r; }
template <int N> WC_HD Dual<N> operator+(float b, const Dual<N>& a) { Dual<N> r = a; r.v += b; return r; }
template <int N> WC_HD Dual<N> operator-(const Dual<N>& a, float b) { Dual<N> r = a; r.v -= b; return r; }
template <int N> WC_HD Dual<N> operator-(float b, const Dual<N>& a) { return make_dual<N>(b) - a; }
template <int N> WC_HD Dual<N> operator*(const Dual<N>& a, float b) {
    Dual<N> r; r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N> WC_HD Dual<N> operator*(float b, const Dual<N>& a) { return a * b; }
template <int N> WC_HD Dual<N> operator/(const Dual<N>& a, float b) { return a * (1.0f / b); }
template <int N> WC_HD Dual<N> operator/(float a, const Dual<N>& b) { return make_dual<N>(a) / b; }
template <int N> WC_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> WC_HD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> WC_HD bool operator<(const Dual<N>& a, float b) { return a.v < b; }
template <int N> WC_HD bool operator>(const Dual<N>& a, float b) { return a.v > b; }

template <int N> WC_HD Dual<N> chain(const Dual<N>& a, float fv, float dfdx) {
    Dual<N> r; r.v = fv;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = dfdx * a.d[i];
    return r;
}

WC_HD float w_sqrt(float x) { return sqrtf(x); }
WC_HD double w_sqrt(double x) { return sqrt(x); }
template <int N> WC_HD Dual<N> w_sqrt(const Dual<N>& a) {
    const float s = sqrtf(a.v);
    return chain(a, s, s > 0.0f ? 0.5f / s : 0.0f);
}
WC_HD float w_abs(float x) { return fabsf(x); }
WC_HD double w_abs(double x) { return fabs(x); }
template <int N> WC_HD Dual<N> w_abs(const Dual<N>& a) { return a.v < 0.0f ? -a : a; }
WC_HD float w_exp(float x) { return expf(x); }
WC_HD double w_exp(double x) { return exp(x); }
template <int N> WC_HD Dual<N> w_exp(const Dual<N>& a) { const float e = expf(a.v); return chain(a, e, e); }
WC_HD float w_sin(float x) { return sinf(x); }
WC_HD double w_sin(double x) { return sin(x); }
template <int N> WC_HD Dual<N> w_sin(const Dual<N>& a) { return chain(a, sinf(a.v), cosf(a.v)); }
WC_HD float w_cos(float x) { return cosf(x); }
WC_HD double w_cos(double x) { return cos(x); }
template <int N> WC_HD Dual<N> w_cos(const Dual<N>& a) { return chain(a, cosf(a.v), -sinf(a.v)); }
WC_HD float w_acos(float x) { return acosf(x); }
WC_HD double w_acos(double x) { return acos(x); }
template <int N> WC_HD Dual<N> w_acos(const Dual<N>& a) {
    const float x = fminf(fmaxf(a.v, -1.0f), 1.0f);
    const float den = sqrtf(fmaxf(1.0f - x * x, 1.0e-12f));
    return chain(a, acosf(x), -1.0f / den);
}
WC_HD float w_atan2(float y, float x) { return atan2f(y, x); }
WC_HD double w_atan2(double y, double x) { return atan2(y, x); }
template <int N> WC_HD Dual<N> w_atan2(const Dual<N>& y, const Dual<N>& x) {
    Dual<N> r; r.v = atan2f(y.v, x.v);
    const float den = x.v * x.v + y.v * y.v;
    const float inv = den > 0.0f ? 1.0f / den : 0.0f;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * inv;
    return r;
}
// max/min/clamp keep the derivative of the selected branch
WC_HD float w_max(float a, float b) { return fmaxf(a, b); }
WC_HD double w_max(double a, double b) { return fmax(a, b); }
template <int N> WC_HD Dual<N> w_max(const Dual<N>& a, float b) { return a.v >= b ? a : make_dual<N>(b); }
WC_HD float w_min(float a, float b) { return fminf(a, b); }
WC_HD double w_min(double a, double b) { return fmin(a, b); }
template <int N> WC_HD Dual<N> w_min(const Dual<N>& a, float b) { return a.v <= b ? a : make_dual<N>(b); }
WC_HD double w_max(double a, float b) { return fmax(a, (double)b); }
WC_HD double w_min(double a, float b) { return fmin(a, (double)b); }

// ---------------------------------------------------------------- 3-vectors
template <typename R>
struct V3 {
    R x, y, z;
};
template <typename R> WC_HD V3<R> v3(R x, R y, R z) { return {x, y, z}; }
template <typename R> WC_HD V3<R> v3c(float x, float y, float z) { return {lift<R>(x), lift<R>(y), lift<R>(z)}; }
template <typename R> WC_HD V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> WC_HD V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> WC_HD V3<R> operator-(const V3<R>& a) { return {-a.x, -a.y, -a.z}; }
template <typename R, typename S> WC_HD V3<R> operator*(const V3<R>& a, const S& s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> WC_HD R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> WC_HD V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> WC_HD R norm(const V3<R>& a) { return w_sqrt(dot(a, a)); }

// Normalize v when |v| > SMALL_EPS, else return the normalized fallback;
// both divide by (norm + EPS) (reference tests/ad/_reference_fields.py:17-22).
template <typename R>
WC_HD V3<R> safe_normalize(const V3<R>& v, const V3<R>& fallback) {
    const R n = norm(v);
    if (val(n) > kSmallEpsF) {
        const R inv = lift<R>(1.0f) / (n + kEpsF);
        return v * inv;
    }
    const R fn = norm(fallback);
    const R inv = lift<R>(1.0f) / (fn + kEpsF);
    return fallback * inv;
}

// Unit vector perpendicular to ray_dir closest to `preferred`; when preferred
// is parallel to the ray, fall back to world z (or y for near-vertical rays)
// (reference tests/ad/_reference_fields.py:25-32).
template <typename R>
WC_HD V3<R> stable_perp_basis(const V3<R>& ray_dir, const V3<R>& preferred) {
    const V3<R> proj = preferred - ray_dir * dot(preferred, ray_dir);
    const V3<R> alt_axis = fabsf(val(ray_dir.z)) < 0.9f ? v3c<R>(0.f, 0.f, 1.f) : v3c<R>(0.f, 1.f, 0.f);
    const V3<R> alt_proj = alt_axis - ray_dir * dot(alt_axis, ray_dir);
    return safe_normalize(proj, alt_proj);
}

// Transverse projection without renormalization: keeps the short-dipole
// sin(theta) weight of a polarization axis.
template <typename R>
WC_HD V3<R> transverse(const V3<R>& ray_dir, const V3<R>& pol) {
    return pol - ray_dir * dot(pol, ray_dir);
}

// ---------------------------------------------------------------- complex
template <typename R>
struct Cx {
    R re, im;
};
template <typename R> WC_HD Cx<R> cx(R re, R im) { return {re, im}; }
template <typename R> WC_HD Cx<R> cxc(float re, float im) { return {lift<R>(re), lift<R>(im)}; }
template <typename R> WC_HD Cx<R> operator+(const Cx<R>& a, const Cx<R>& b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> WC_HD Cx<R> operator-(const Cx<R>& a, const Cx<R>& b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> WC_HD Cx<R> operator-(const Cx<R>& a) { return {-a.re, -a.im}; }
template <typename R> WC_HD Cx<R> operator*(const Cx<R>& a, const Cx<R>& b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R, typename S> WC_HD Cx<R> scale(const Cx<R>& a, const S& s) { return {a.re * s, a.im * s}; }
template <typename R> WC_HD Cx<R> operator/(const Cx<R>& a, const Cx<R>& b) {
    const R den = b.re * b.re + b.im * b.im;
    const R inv = lift<R>(1.0f) / den;
    return {(a.re * b.re + a.im * b.im) * inv, (a.im * b.re - a.re * b.im) * inv};
}
template <typename R> WC_HD R abs2(const Cx<R>& a) { return a.re * a.re + a.im * a.im; }

// Principal square root (branch cut on the negative real axis), written to
// avoid cancellation: the larger component comes from sqrt((|z| +- x) / 2)
// and the other from y / (2 * larger).
template <typename R>
WC_HD Cx<R> csqrt(const Cx<R>& z) {
    const R mag = w_sqrt(z.re * z.re + z.im * z.im);
    if (val(mag) == 0.0f) return cxc<R>(0.f, 0.f);
    if (val(z.re) >= 0.0f) {
        const R re = w_sqrt((mag + z.re) * 0.5f);
        return {re, z.im / (re * 2.0f)};
    }
    R im = w_sqrt((mag - z.re) * 0.5f);
    if (val(z.im) < 0.0f) im = -im;
    return {z.im / (im * 2.0f), im};
}
template <typename R>
WC_HD Cx<R> cexp(const Cx<R>& z) {
    const R m = w_exp(z.re);
    return {m * w_cos(z.im), m * w_sin(z.im)};
}
// e^{j*phi}
template <typename R>
WC_HD Cx<R> cexpj(const R& phi) { return {w_cos(phi), w_sin(phi)}; }

template <typename R>
struct C3 {
    Cx<R> x, y, z;
};
template <typename R> WC_HD C3<R> c3_from_real(const V3<R>& v) {
    return {{v.x, lift<R>(0.f)}, {v.y, lift<R>(0.f)}, {v.z, lift<R>(0.f)}};
}
template <typename R> WC_HD Cx<R> c3_dot_real(const C3<R>& a, const V3<R>& v) {
    return {a.x.re * v.x + a.y.re * v.y + a.z.re * v.z, a.x.im * v.x + a.y.im * v.y + a.z.im * v.z};
}
template <typename R> WC_HD C3<R> c3_axis(const V3<R>& axis, const Cx<R>& s) {
    return {scale(s, axis.x), scale(s, axis.y), scale(s, axis.z)};
}
template <typename R> WC_HD C3<R> operator+(const C3<R>& a, const C3<R>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> WC_HD C3<R> c3_scale(const C3<R>& a, const Cx<R>& s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> WC_HD R c3_power(const C3<R>& a) { return abs2(a.x) + abs2(a.y) + abs2(a.z); }

// -k*d reduced modulo 2*pi with the product formed in double precision: at
// mmWave and km ranges k*d reaches 1e6 rad and a float product would lose
// ~0.06 rad of phase.
WC_HD float neg_kd_phase(double k, double d) {
    double p = -k * d;
    p = p - kTwoPi * floor(p / kTwoPi);
    return (float)p;
}

}  // namespace wc
