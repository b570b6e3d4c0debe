// Row-level field operators with their tangent (JVP) and adjoint (VJP)
// companions (reference witwin/channel/kernels/fields.py:110-830,
// kernels/materials.py:151-262; the native ops live in RayD and
// native/channel/kernels/fields.cu:36-204, 1237, 1873, 2183).
//
// Every operator is one templated row function evaluated in float (the
// forward), in Dual<1> seeded with the input tangents (JVP) or in Dual<8>
// passes seeded one-hot over the row's own input scalars (VJP: each pass
// contracts the output cotangents with eight Jacobian columns). Inputs shared
// by rows (the frequency, layer tables) get their gradients from per-row
// partials reduced on the device.
#include "wc_api_internal.h"
#include "wc_paths.h"
#include "wc_tangent.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

constexpr int kRowLayers = 8;  // layers of one material a row can differentiate
using R8 = Dual<8>;
using R1 = Dual<1>;

template <typename R> struct RowOut {
    C3<R> field;
    Cx<R> coef, pf;
    R gain, length, delay;
    V3<R> dir;
};

template <typename R>
__device__ __forceinline__ void finish_row(RowOut<R>& o, const V3<float>& rx_pol, float sqrt_power) {
    o.coef = project_receiver(o.field, o.dir, lift3<R>(rx_pol));
    o.pf = scale(o.coef, lift<R>(sqrt_power));
    o.gain = o.pf.re * o.pf.re + o.pf.im * o.pf.im;
    o.delay = o.length * (float)(1.0 / kC0);
}

// free space (fields.cu:36-82): E = p_tx,perp / (2 k L) e^{-jkL}
template <typename R>
__device__ void free_space_row(const V3<R>& a, const V3<R>& b, const V3<float>& tx_pol, const V3<float>& rx_pol,
                               float sqrt_power, const R& freq, double freq_d, RowOut<R>& o) {
    const V3<R> seg = b - a;
    const R L = norm(seg);
    const double L_d = norm(to_d(vval(b)) - to_d(vval(a)));
    const double k_d = freq_d * kTwoPi / kC0;
    o.dir = safe_normalize(seg, v3c<R>(0.f, 0.f, 1.f));
    const R k = wave_number(freq);
    const R amp = lift<R>(1.0f) / (w_max(k, kSmallEpsF) * w_max(L, kEpsF) * 2.0f);
    o.field = c3_axis(transverse(o.dir, lift3<R>(tx_pol)), scale(cexpj(PhaseOf<R>::make(k, L, k_d, L_d)), amp));
    o.length = L;
    finish_row(o, rx_pol, sqrt_power);
}

// reflection sequence (fields.cu:113-204): Jones transport through D slab
// reflections, amplitude 1/(2 k L), phase -k L with L the unfolded length
template <typename R>
__device__ void reflection_row(int depth, const V3<R>& src, const V3<R>& tgt, const V3<R>* pos, const V3<R>* nrm,
                               const R* eps, const R* sig, const float* mu, const R* gain, const R* thick,
                               const V3<float>& tx_pol, const V3<float>& rx_pol, float sqrt_power, const R& freq,
                               double freq_d, RowOut<R>& o) {
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    V3<R> prev = src;
    V3<R> incident = safe_normalize(pos[0] - prev, ez);
    C3<R> value = c3_from_real(transverse(incident, lift3<R>(tx_pol)));
    V3<R> outgoing = incident;
    R total = lift<R>(0.0f);
    double total_d = 0.0;
    for (int b = 0; b < depth; ++b) {
        incident = safe_normalize(pos[b] - prev, outgoing);
        total = total + norm(pos[b] - prev);
        total_d += norm(to_d(vval(pos[b])) - to_d(vval(prev)));
        value = reflect_c3(value, incident, nrm[b], eps[b], sig[b], lift<R>(mu[b]), gain[b], thick[b], freq,
                           outgoing);
        prev = pos[b];
    }
    o.dir = safe_normalize(tgt - prev, outgoing);
    total = total + norm(tgt - prev);
    total_d += norm(to_d(vval(tgt)) - to_d(vval(prev)));
    const double k_d = freq_d * kTwoPi / kC0;
    const R k = wave_number(freq);
    const R amp = lift<R>(1.0f) / (w_max(k, kSmallEpsF) * w_max(total, kEpsF) * 2.0f);
    o.field = c3_scale(value, scale(cexpj(PhaseOf<R>::make(k, total, k_d, total_d)), amp));
    o.length = total;
    finish_row(o, rx_pol, sqrt_power);
}

// Layer view of one interaction: the material's layers with optional R
// overrides of thickness / eps_r / sigma_e (kRowLayers at most).
struct LayerTables {
    const int* offset;
    const int* count;
    const float* thickness;
    const float* eps;
    const float* sigma;
    const float* mu;
    int materials;
    int total;
};

// material m names a layer span inside the tables
__device__ __forceinline__ bool material_ok(const LayerTables& lt, int m) {
    if (m < 0 || m >= lt.materials) return false;
    const int lo = lt.offset[m], lc = lt.count[m];
    return lo >= 0 && lc >= 0 && lc <= kRowLayers && lo + lc <= lt.total;
}

// straight transmission through thin layered walls (reference
// tests/ad/_reference_fields.py:269-336): per valid interaction the Rouard
// t_TE / t_TM on the wall's s/p basis; carrier L - sum(thickness cos);
// amplitude 1/(2 k L) with L the full straight length. layer(k, i, which,
// base) gives the R value of table entry i (which: 0 thickness, 1 eps_r,
// 2 sigma_e) at interaction k, so each caller seeds its own tangents.
template <typename R, typename LayerValue>
__device__ void transmission_row(int depth, const unsigned char* ivalid, const int* mat, const V3<R>& src,
                                 const V3<R>& tgt, const V3<R>* nrm, const LayerTables& lt, LayerValue layer,
                                 const V3<float>& tx_pol, const V3<float>& rx_pol, float sqrt_power, const R& freq,
                                 double freq_d, RowOut<R>& o) {
    const V3<R> seg = tgt - src;
    const R L = norm(seg);
    o.dir = safe_normalize(seg, v3c<R>(0.f, 0.f, 1.f));
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    C3<R> value = c3_from_real(transverse(o.dir, lift3<R>(tx_pol)));
    R carrier = L;
    double carrier_d = norm(to_d(vval(tgt)) - to_d(vval(src)));
    for (int k = 0; k < depth; ++k) {
        if (!ivalid[k]) continue;
        V3<R> n = safe_normalize(nrm[k], ez);
        if (val(dot(o.dir, n)) > 0.0f) n = -n;
        const R ct = w_min(w_max(w_abs(dot(o.dir, n)), kSmallEpsF), 1.0f);
        const int lo = lt.offset[mat[k]];
        const int lc = min(lt.count[mat[k]], kRowLayers);
        R th[kRowLayers], ep[kRowLayers], sg[kRowLayers];
        R wall = lift<R>(0.0f);
        for (int l = 0; l < lc; ++l) {
            th[l] = layer(k, lo + l, 0, lt.thickness[lo + l]);
            ep[l] = layer(k, lo + l, 1, lt.eps[lo + l]);
            sg[l] = layer(k, lo + l, 2, lt.sigma[lo + l]);
            wall = wall + w_max(th[l], 0.0f);
        }
        Cx<R> r;
        const Cx<R> t_te = stack_rt<R>(ct, lt.thickness + lo, lt.eps + lo, lt.sigma + lo, lt.mu + lo, lc, freq, 0,
                                       ep, sg, th, r);
        const Cx<R> t_tm = stack_rt<R>(ct, lt.thickness + lo, lt.eps + lo, lt.sigma + lo, lt.mu + lo, lc, freq, 1,
                                       ep, sg, th, r);
        const V3<R> sax = safe_normalize(cross(n, o.dir), stable_perp_basis(o.dir, n));
        const V3<R> pax = safe_normalize(cross(sax, o.dir), stable_perp_basis(o.dir, sax));
        const Cx<R> e_s = c3_dot_real(value, sax);
        const Cx<R> e_p = c3_dot_real(value, pax);
        value = c3_axis(sax, t_te * e_s) + c3_axis(pax, t_tm * e_p);
        carrier = carrier - ct * wall;
        carrier_d -= (double)val(ct) * (double)val(wall);
    }
    const double k_d = freq_d * kTwoPi / kC0;
    const R k = wave_number(freq);
    const R amp = lift<R>(1.0f) / (w_max(k, kSmallEpsF) * w_max(L, kEpsF) * 2.0f);
    o.field = c3_scale(value, scale(cexpj(PhaseOf<R>::make(k, carrier, k_d, carrier_d)), amp));
    o.length = L;
    finish_row(o, rx_pol, sqrt_power);
}

// layer stack between vacuum half-spaces (reference kernels/materials.py:151):
// r, t per polarization and the power coefficients R = |r|^2, T = |t|^2 (both
// bounding media are vacuum, so the flux factor Re(Y2)/Re(Y1) is 1)
template <typename R> struct StackOut {
    Cx<R> r_te, r_tm, t_te, t_tm;
    R R_te, R_tm, T_te, T_tm;
};

template <typename R>
__device__ void stack_row(const R& cos_theta, int m, const LayerTables& lt, const R* thick, const R* eps,
                          const R* sig, const R& freq, StackOut<R>& o) {
    const int lo = lt.offset[m];
    const int lc = min(lt.count[m], kRowLayers);
    o.t_te = stack_rt<R>(cos_theta, lt.thickness + lo, lt.eps + lo, lt.sigma + lo, lt.mu + lo, lc, freq, 0, eps, sig,
                         thick, o.r_te);
    o.t_tm = stack_rt<R>(cos_theta, lt.thickness + lo, lt.eps + lo, lt.sigma + lo, lt.mu + lo, lc, freq, 1, eps, sig,
                         thick, o.r_tm);
    o.R_te = o.r_te.re * o.r_te.re + o.r_te.im * o.r_te.im;
    o.R_tm = o.r_tm.re * o.r_tm.re + o.r_tm.im * o.r_tm.im;
    o.T_te = o.t_te.re * o.t_te.re + o.t_te.im * o.t_te.im;
    o.T_tm = o.t_tm.re * o.t_tm.re + o.t_tm.im * o.t_tm.im;
}

// ------------------------------------------------------------------ seeding
template <typename R> __device__ __forceinline__ R jv(float v, float t);
template <> __device__ __forceinline__ R1 jv<R1>(float v, float t) {
    R1 r = make_dual<1>(v);
    r.d[0] = t;
    return r;
}
// one-hot seed of local slot `slot` in pass `pass` of Dual<8>
__device__ __forceinline__ R8 sv(float v, int slot, int pass) {
    const int local = slot - 8 * pass;
    return seed<8>(v, (local >= 0 && local < 8) ? local : -1);
}

__device__ __forceinline__ V3<float> ld3o(const float* p, size_t i) {
    return p ? ld3(p + 3 * i) : v3c<float>(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float ldo(const float* p, size_t i) { return p ? p[i] : 0.0f; }

// Output views and cotangents of the shared enumerated-field schema.
struct FieldOutPtrs {
    float* field_vector;  // [N, 3] c64
    float* coefficient;   // [N] c64
    float* path_field;    // [N] c64
    float* path_gain;
    float* path_length;
    float* delay;
    float* direction;  // [N, 3]
};

template <typename R>
__device__ __forceinline__ void store_value(const FieldOutPtrs& o, size_t i, const RowOut<R>& r) {
    const C3<R>& f = r.field;
    float* fv = o.field_vector + 6 * i;
    fv[0] = val(f.x.re); fv[1] = val(f.x.im); fv[2] = val(f.y.re); fv[3] = val(f.y.im);
    fv[4] = val(f.z.re); fv[5] = val(f.z.im);
    o.coefficient[2 * i] = val(r.coef.re);
    o.coefficient[2 * i + 1] = val(r.coef.im);
    o.path_field[2 * i] = val(r.pf.re);
    o.path_field[2 * i + 1] = val(r.pf.im);
    o.path_gain[i] = val(r.gain);
    o.path_length[i] = val(r.length);
    o.delay[i] = val(r.delay);
    if (o.direction) {
        o.direction[3 * i] = val(r.dir.x);
        o.direction[3 * i + 1] = val(r.dir.y);
        o.direction[3 * i + 2] = val(r.dir.z);
    }
}

__device__ __forceinline__ void store_tangent(const FieldOutPtrs& o, size_t i, const RowOut<R1>& r) {
    const C3<R1>& f = r.field;
    float* fv = o.field_vector + 6 * i;
    fv[0] = f.x.re.d[0]; fv[1] = f.x.im.d[0]; fv[2] = f.y.re.d[0]; fv[3] = f.y.im.d[0];
    fv[4] = f.z.re.d[0]; fv[5] = f.z.im.d[0];
    o.coefficient[2 * i] = r.coef.re.d[0];
    o.coefficient[2 * i + 1] = r.coef.im.d[0];
    o.path_field[2 * i] = r.pf.re.d[0];
    o.path_field[2 * i + 1] = r.pf.im.d[0];
    o.path_gain[i] = r.gain.d[0];
    o.path_length[i] = r.length.d[0];
    o.delay[i] = r.delay.d[0];
    if (o.direction) {
        o.direction[3 * i] = r.dir.x.d[0];
        o.direction[3 * i + 1] = r.dir.y.d[0];
        o.direction[3 * i + 2] = r.dir.z.d[0];
    }
}

struct FieldGrads {
    const float* field_vector;
    const float* coefficient;
    const float* path_field;
    const float* path_gain;
    const float* path_length;
    const float* delay;
    const float* direction;
};

// sum over outputs of <cotangent, d output / d slot j> for the 8 slots of a pass
__device__ __forceinline__ void contract_row(const FieldGrads& g, size_t i, const RowOut<R8>& r, float d[8]) {
    for (int j = 0; j < 8; ++j) d[j] = 0.0f;
    auto add = [&](const float* gp, size_t k, const R8& x) {
        if (!gp) return;
        const float w = gp[k];
        if (w == 0.0f) return;
#pragma unroll
        for (int j = 0; j < 8; ++j) d[j] += w * x.d[j];
    };
    const C3<R8>& f = r.field;
    add(g.field_vector, 6 * i, f.x.re); add(g.field_vector, 6 * i + 1, f.x.im);
    add(g.field_vector, 6 * i + 2, f.y.re); add(g.field_vector, 6 * i + 3, f.y.im);
    add(g.field_vector, 6 * i + 4, f.z.re); add(g.field_vector, 6 * i + 5, f.z.im);
    add(g.coefficient, 2 * i, r.coef.re); add(g.coefficient, 2 * i + 1, r.coef.im);
    add(g.path_field, 2 * i, r.pf.re); add(g.path_field, 2 * i + 1, r.pf.im);
    add(g.path_gain, i, r.gain);
    add(g.path_length, i, r.length);
    add(g.delay, i, r.delay);
    add(g.direction, 3 * i, r.dir.x); add(g.direction, 3 * i + 1, r.dir.y); add(g.direction, 3 * i + 2, r.dir.z);
}

// ------------------------------------------------------------------ free space
__global__ void __launch_bounds__(128) fs_jvp_kernel(int n, const float* __restrict__ src, const float* __restrict__ tgt,
                              const float* __restrict__ power, const float* __restrict__ txp,
                              const float* __restrict__ rxp, double freq, float t_freq, const float* __restrict__ t_src,
                              const float* __restrict__ t_tgt, FieldOutPtrs o) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const V3<float> a = ld3(src + 3 * i), b = ld3(tgt + 3 * i), ta = ld3o(t_src, i), tb = ld3o(t_tgt, i);
        const V3<R1> A{jv<R1>(a.x, ta.x), jv<R1>(a.y, ta.y), jv<R1>(a.z, ta.z)};
        const V3<R1> B{jv<R1>(b.x, tb.x), jv<R1>(b.y, tb.y), jv<R1>(b.z, tb.z)};
        RowOut<R1> r;
        free_space_row<R1>(A, B, ld3(txp + 3 * i), ld3(rxp + 3 * i), sqrtf(fmaxf(power[i], 0.0f)),
                           jv<R1>((float)freq, t_freq), freq, r);
        store_tangent(o, i, r);
    }
}

// slots: 0-2 source, 3-5 target, 6 frequency
__global__ void __launch_bounds__(128) fs_vjp_kernel(int n, const float* __restrict__ src, const float* __restrict__ tgt,
                              const float* __restrict__ power, const float* __restrict__ txp,
                              const float* __restrict__ rxp, double freq, FieldGrads g, float* __restrict__ g_src,
                              float* __restrict__ g_tgt, float* __restrict__ g_freq_row) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const V3<float> a = ld3(src + 3 * i), b = ld3(tgt + 3 * i);
        const V3<R8> A{sv(a.x, 0, 0), sv(a.y, 1, 0), sv(a.z, 2, 0)};
        const V3<R8> B{sv(b.x, 3, 0), sv(b.y, 4, 0), sv(b.z, 5, 0)};
        RowOut<R8> r;
        free_space_row<R8>(A, B, ld3(txp + 3 * i), ld3(rxp + 3 * i), sqrtf(fmaxf(power[i], 0.0f)),
                           sv((float)freq, 6, 0), freq, r);
        float d[8];
        contract_row(g, i, r, d);
        if (g_src) { g_src[3 * i] = d[0]; g_src[3 * i + 1] = d[1]; g_src[3 * i + 2] = d[2]; }
        if (g_tgt) { g_tgt[3 * i] = d[3]; g_tgt[3 * i + 1] = d[4]; g_tgt[3 * i + 2] = d[5]; }
        g_freq_row[i] = d[6];
    }
}

// ------------------------------------------------------------------ reflection sequence
struct ReflIn {
    int depth;
    const float *src, *tgt, *pos, *nrm, *power, *txp, *rxp, *eps, *sig, *mu, *gain, *thick;
};

__global__ void __launch_bounds__(128) refl_jvp_kernel(int n, ReflIn in, double freq, float t_freq, ReflIn t, FieldOutPtrs o) {
    const int D = in.depth;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        auto v3 = [&](const float* p, const float* tp, size_t k) {
            const V3<float> a = ld3(p + 3 * k), b = ld3o(tp, k);
            return V3<R1>{jv<R1>(a.x, b.x), jv<R1>(a.y, b.y), jv<R1>(a.z, b.z)};
        };
        V3<R1> pos[kMaxDepth], nrm[kMaxDepth];
        R1 eps[kMaxDepth], sig[kMaxDepth], gain[kMaxDepth], thick[kMaxDepth];
        float mu[kMaxDepth];
        for (int b = 0; b < D; ++b) {
            const size_t k = (size_t)i * D + b;
            pos[b] = v3(in.pos, t.pos, k);
            nrm[b] = v3(in.nrm, t.nrm, k);
            eps[b] = jv<R1>(in.eps[k], ldo(t.eps, k));
            sig[b] = jv<R1>(in.sig[k], ldo(t.sig, k));
            gain[b] = jv<R1>(in.gain[k], ldo(t.gain, k));
            thick[b] = jv<R1>(in.thick[k], ldo(t.thick, k));
            mu[b] = in.mu[k];
        }
        RowOut<R1> r;
        reflection_row<R1>(D, v3(in.src, t.src, i), v3(in.tgt, t.tgt, i), pos, nrm, eps, sig, mu, gain, thick,
                           ld3(in.txp + 3 * i), ld3(in.rxp + 3 * i), sqrtf(fmaxf(in.power[i], 0.0f)),
                           jv<R1>((float)freq, t_freq), freq, r);
        store_tangent(o, i, r);
    }
}

// slots: 0-2 source, 3-5 target, 6 frequency, then per bounce b (base 7 + 10 b):
// position xyz, normal xyz, eps, sigma, gain, thickness
struct ReflGrads {
    float *src, *tgt, *pos, *nrm, *eps, *sig, *gain, *thick, *freq_row;
};

__global__ void __launch_bounds__(128) refl_vjp_kernel(int n, ReflIn in, double freq, FieldGrads g, ReflGrads out) {
    const int D = in.depth;
    const int slots = 7 + 10 * D;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        float fr = 0.0f;
        for (int pass = 0; pass * 8 < slots; ++pass) {
            auto v3 = [&](const float* p, size_t k, int s0) {
                const V3<float> a = ld3(p + 3 * k);
                return V3<R8>{sv(a.x, s0, pass), sv(a.y, s0 + 1, pass), sv(a.z, s0 + 2, pass)};
            };
            V3<R8> pos[kMaxDepth], nrm[kMaxDepth];
            R8 eps[kMaxDepth], sig[kMaxDepth], gain[kMaxDepth], thick[kMaxDepth];
            float mu[kMaxDepth];
            for (int b = 0; b < D; ++b) {
                const size_t k = (size_t)i * D + b;
                const int s = 7 + 10 * b;
                pos[b] = v3(in.pos, k, s);
                nrm[b] = v3(in.nrm, k, s + 3);
                eps[b] = sv(in.eps[k], s + 6, pass);
                sig[b] = sv(in.sig[k], s + 7, pass);
                gain[b] = sv(in.gain[k], s + 8, pass);
                thick[b] = sv(in.thick[k], s + 9, pass);
                mu[b] = in.mu[k];
            }
            RowOut<R8> r;
            reflection_row<R8>(D, v3(in.src, i, 0), v3(in.tgt, i, 3), pos, nrm, eps, sig, mu, gain, thick,
                               ld3(in.txp + 3 * i), ld3(in.rxp + 3 * i), sqrtf(fmaxf(in.power[i], 0.0f)),
                               sv((float)freq, 6, pass), freq, r);
            float d[8];
            contract_row(g, i, r, d);
            for (int j = 0; j < 8; ++j) {
                const int s = 8 * pass + j;
                if (s >= slots) break;
                if (s < 3) { if (out.src) out.src[3 * i + s] = d[j]; }
                else if (s < 6) { if (out.tgt) out.tgt[3 * i + s - 3] = d[j]; }
                else if (s == 6) fr = d[j];
                else {
                    const int b = (s - 7) / 10, q = (s - 7) % 10;
                    const size_t k = (size_t)i * D + b;
                    if (q < 3) { if (out.pos) out.pos[3 * k + q] = d[j]; }
                    else if (q < 6) { if (out.nrm) out.nrm[3 * k + q - 3] = d[j]; }
                    else if (q == 6) { if (out.eps) out.eps[k] = d[j]; }
                    else if (q == 7) { if (out.sig) out.sig[k] = d[j]; }
                    else if (q == 8) { if (out.gain) out.gain[k] = d[j]; }
                    else { if (out.thick) out.thick[k] = d[j]; }
                }
            }
        }
        out.freq_row[i] = fr;
    }
}

// ------------------------------------------------------------------ transmission sequence
struct TransIn {
    int depth;
    const unsigned char* path_valid;
    const unsigned char* ivalid;
    const int* mat;
    const float *src, *tgt, *nrm, *power, *txp, *rxp;
    LayerTables lt;
};

// A row is evaluated when the path is valid and every valid interaction names
// a material of the tables; otherwise its outputs (and gradients) are zero.
__device__ __forceinline__ bool trans_row_ok(const TransIn& in, int i) {
    if (!in.path_valid[i]) return false;
    for (int k = 0; k < in.depth; ++k) {
        const size_t q = (size_t)i * in.depth + k;
        if (in.ivalid[q] && !material_ok(in.lt, in.mat[q])) return false;
    }
    return true;
}

__device__ __forceinline__ void zero_row(const FieldOutPtrs& o, size_t i) {
    for (int j = 0; j < 6; ++j) o.field_vector[6 * i + j] = 0.0f;
    o.coefficient[2 * i] = o.coefficient[2 * i + 1] = 0.0f;
    o.path_field[2 * i] = o.path_field[2 * i + 1] = 0.0f;
    o.path_gain[i] = o.path_length[i] = o.delay[i] = 0.0f;
    if (o.direction) o.direction[3 * i] = o.direction[3 * i + 1] = o.direction[3 * i + 2] = 0.0f;
}

// interaction flags, materials and normals of row i
template <typename R, typename NormalValue>
__device__ __forceinline__ void trans_load(const TransIn& in, int i, unsigned char* iv, int* mat, V3<R>* nrm,
                                           NormalValue normal) {
    for (int k = 0; k < in.depth; ++k) {
        const size_t q = (size_t)i * in.depth + k;
        iv[k] = in.ivalid[q];
        mat[k] = iv[k] ? in.mat[q] : 0;
        nrm[k] = normal(k, q, ld3(in.nrm + 3 * q));
    }
}

__global__ void __launch_bounds__(128) trans_fwd_kernel(int n, TransIn in, double freq, FieldOutPtrs o) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        if (!trans_row_ok(in, i)) {
            zero_row(o, i);
            continue;
        }
        unsigned char iv[kMaxHits];
        int mat[kMaxHits];
        V3<float> nrm[kMaxHits];
        trans_load<float>(in, i, iv, mat, nrm, [](int, size_t, const V3<float>& a) { return a; });
        RowOut<float> r;
        transmission_row<float>(in.depth, iv, mat, ld3(in.src + 3 * i), ld3(in.tgt + 3 * i), nrm, in.lt,
                                [](int, int, int, float base) { return base; }, ld3(in.txp + 3 * i),
                                ld3(in.rxp + 3 * i), sqrtf(fmaxf(in.power[i], 0.0f)), (float)freq, freq, r);
        store_value(o, i, r);
    }
}

struct TransTangents {
    const float *src, *tgt, *nrm, *thick, *eps, *sig;  // layer tangents index the layer tables
};

__global__ void __launch_bounds__(128) trans_jvp_kernel(int n, TransIn in, double freq, float t_freq, TransTangents t, FieldOutPtrs o) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        if (!trans_row_ok(in, i)) {
            zero_row(o, i);
            continue;
        }
        unsigned char iv[kMaxHits];
        int mat[kMaxHits];
        V3<R1> nrm[kMaxHits];
        trans_load<R1>(in, i, iv, mat, nrm, [&](int, size_t q, const V3<float>& a) {
            const V3<float> b = ld3o(t.nrm, q);
            return V3<R1>{jv<R1>(a.x, b.x), jv<R1>(a.y, b.y), jv<R1>(a.z, b.z)};
        });
        const V3<float> a = ld3(in.src + 3 * i), b = ld3(in.tgt + 3 * i), ta = ld3o(t.src, i), tb = ld3o(t.tgt, i);
        RowOut<R1> r;
        transmission_row<R1>(
            in.depth, iv, mat, {jv<R1>(a.x, ta.x), jv<R1>(a.y, ta.y), jv<R1>(a.z, ta.z)},
            {jv<R1>(b.x, tb.x), jv<R1>(b.y, tb.y), jv<R1>(b.z, tb.z)}, nrm, in.lt,
            [&](int, int e, int which, float base) {
                const float* tp = which == 0 ? t.thick : which == 1 ? t.eps : t.sig;
                return jv<R1>(base, ldo(tp, e));
            },
            ld3(in.txp + 3 * i), ld3(in.rxp + 3 * i), sqrtf(fmaxf(in.power[i], 0.0f)), jv<R1>((float)freq, t_freq),
            freq, r);
        store_tangent(o, i, r);
    }
}

// Gradient slots of one row: 0-2 source, 3-5 target, 6 frequency, then per
// interaction k from base[k]: normal xyz and, when the interaction is valid,
// thickness / eps_r / sigma_e of each of its layers. Layer gradients of all
// rows meet in float64 per-layer sums.
struct TransGrads {
    float *src, *tgt, *nrm, *freq_row;
    double *thick, *eps, *sig;  // [layer_total]
};

__global__ void __launch_bounds__(128) trans_vjp_kernel(int n, TransIn in, double freq, FieldGrads g, TransGrads out) {
    const int D = in.depth;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        if (out.src) out.src[3 * i] = out.src[3 * i + 1] = out.src[3 * i + 2] = 0.0f;
        if (out.tgt) out.tgt[3 * i] = out.tgt[3 * i + 1] = out.tgt[3 * i + 2] = 0.0f;
        if (out.nrm)
            for (int j = 0; j < 3 * D; ++j) out.nrm[3 * (size_t)i * D + j] = 0.0f;
        out.freq_row[i] = 0.0f;
        if (!trans_row_ok(in, i)) continue;
        int base[kMaxHits + 1];
        base[0] = 7;
        for (int k = 0; k < D; ++k) {
            const size_t q = (size_t)i * D + k;
            const int lc = in.ivalid[q] ? min(in.lt.count[in.mat[q]], kRowLayers) : 0;
            base[k + 1] = base[k] + 3 + 3 * lc;
        }
        const int slots = base[D];
        for (int pass = 0; pass * 8 < slots; ++pass) {
            unsigned char iv[kMaxHits];
            int mat[kMaxHits];
            V3<R8> nrm[kMaxHits];
            trans_load<R8>(in, i, iv, mat, nrm, [&](int k, size_t, const V3<float>& a) {
                return V3<R8>{sv(a.x, base[k], pass), sv(a.y, base[k] + 1, pass), sv(a.z, base[k] + 2, pass)};
            });
            const V3<float> a = ld3(in.src + 3 * i), b = ld3(in.tgt + 3 * i);
            RowOut<R8> r;
            transmission_row<R8>(
                D, iv, mat, {sv(a.x, 0, pass), sv(a.y, 1, pass), sv(a.z, 2, pass)},
                {sv(b.x, 3, pass), sv(b.y, 4, pass), sv(b.z, 5, pass)}, nrm, in.lt,
                [&](int k, int e, int which, float v) {
                    return sv(v, base[k] + 3 + 3 * (e - in.lt.offset[mat[k]]) + which, pass);
                },
                ld3(in.txp + 3 * i), ld3(in.rxp + 3 * i), sqrtf(fmaxf(in.power[i], 0.0f)), sv((float)freq, 6, pass),
                freq, r);
            float d[8];
            contract_row(g, i, r, d);
            for (int j = 0; j < 8; ++j) {
                const int s = 8 * pass + j;
                if (s >= slots) break;
                if (s < 3) {
                    if (out.src) out.src[3 * i + s] = d[j];
                } else if (s < 6) {
                    if (out.tgt) out.tgt[3 * i + s - 3] = d[j];
                } else if (s == 6) {
                    out.freq_row[i] = d[j];
                } else {
                    int k = 0;
                    while (s >= base[k + 1]) ++k;
                    const int qq = s - base[k];
                    const size_t q = (size_t)i * D + k;
                    if (qq < 3) {
                        if (out.nrm) out.nrm[3 * q + qq] = d[j];
                    } else if (d[j] != 0.0f) {
                        const int l = (qq - 3) / 3, which = (qq - 3) % 3;
                        double* dst = which == 0 ? out.thick : which == 1 ? out.eps : out.sig;
                        if (dst) unsafeAtomicAdd(dst + in.lt.offset[mat[k]] + l, (double)d[j]);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------------ layer stack
struct StackIn {
    const float* cos_theta;
    const int* mat;
    LayerTables lt;
};
struct StackPtrs {
    float *r_te_re, *r_te_im, *r_tm_re, *r_tm_im, *t_te_re, *t_te_im, *t_tm_re, *t_tm_im, *R_te, *R_tm, *T_te, *T_tm;
};

template <typename R, typename Pick>
__device__ __forceinline__ void stack_store(const StackPtrs& o, int i, const StackOut<R>& s, Pick pick) {
    o.r_te_re[i] = pick(s.r_te.re); o.r_te_im[i] = pick(s.r_te.im);
    o.r_tm_re[i] = pick(s.r_tm.re); o.r_tm_im[i] = pick(s.r_tm.im);
    o.t_te_re[i] = pick(s.t_te.re); o.t_te_im[i] = pick(s.t_te.im);
    o.t_tm_re[i] = pick(s.t_tm.re); o.t_tm_im[i] = pick(s.t_tm.im);
    o.R_te[i] = pick(s.R_te); o.R_tm[i] = pick(s.R_tm); o.T_te[i] = pick(s.T_te); o.T_tm[i] = pick(s.T_tm);
}

__device__ __forceinline__ bool stack_row_ok(const StackIn& in, int i) {
    return material_ok(in.lt, in.mat[i]);
}

// A row whose material is outside the tables gets zero outputs.
__device__ __forceinline__ void stack_zero(const StackPtrs& o, int i) {
    o.r_te_re[i] = o.r_te_im[i] = o.r_tm_re[i] = o.r_tm_im[i] = 0.0f;
    o.t_te_re[i] = o.t_te_im[i] = o.t_tm_re[i] = o.t_tm_im[i] = 0.0f;
    o.R_te[i] = o.R_tm[i] = o.T_te[i] = o.T_tm[i] = 0.0f;
}

__global__ void __launch_bounds__(128) stack_fwd_kernel(int n, StackIn in, double freq, StackPtrs o) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        if (!stack_row_ok(in, i)) {
            stack_zero(o, i);
            continue;
        }
        const int m = in.mat[i];
        StackOut<float> s;
        stack_row<float>(in.cos_theta[i], m, in.lt, nullptr, nullptr, nullptr, (float)freq, s);
        stack_store(o, i, s, [](float x) { return x; });
    }
}

struct StackTangents {
    const float *cos_theta, *thick, *eps, *sig;
};

__global__ void __launch_bounds__(128) stack_jvp_kernel(int n, StackIn in, double freq, float t_freq, StackTangents t, StackPtrs o) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        if (!stack_row_ok(in, i)) {
            stack_zero(o, i);
            continue;
        }
        const int m = in.mat[i];
        const int lo = in.lt.offset[m], lc = min(in.lt.count[m], kRowLayers);
        R1 thd[kRowLayers], epd[kRowLayers], sgd[kRowLayers];
        for (int l = 0; l < lc; ++l) {
            thd[l] = jv<R1>(in.lt.thickness[lo + l], ldo(t.thick, lo + l));
            epd[l] = jv<R1>(in.lt.eps[lo + l], ldo(t.eps, lo + l));
            sgd[l] = jv<R1>(in.lt.sigma[lo + l], ldo(t.sig, lo + l));
        }
        StackOut<R1> s;
        stack_row<R1>(jv<R1>(in.cos_theta[i], ldo(t.cos_theta, i)), m, in.lt, thd, epd, sgd,
                      jv<R1>((float)freq, t_freq), s);
        stack_store(o, i, s, [](const R1& x) { return x.d[0]; });
    }
}

struct StackGradIn {
    const float *r_te_re, *r_te_im, *r_tm_re, *r_tm_im, *t_te_re, *t_te_im, *t_tm_re, *t_tm_im, *R_te, *R_tm, *T_te,
        *T_tm;
};

// slots: 0 cos_theta, 1 frequency, then per layer l (2 + 3 l): thickness,
// eps_r, sigma_e
__global__ void __launch_bounds__(128) stack_vjp_kernel(int n, StackIn in, double freq, StackGradIn g, float* __restrict__ g_cos,
                                 float* __restrict__ g_freq_row, double* __restrict__ g_thick,
                                 double* __restrict__ g_eps, double* __restrict__ g_sig) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        if (!stack_row_ok(in, i)) {
            if (g_cos) g_cos[i] = 0.0f;
            g_freq_row[i] = 0.0f;
            continue;
        }
        const int m = in.mat[i];
        const int lo = in.lt.offset[m], lc = min(in.lt.count[m], kRowLayers);
        const int slots = 2 + 3 * lc;
        float gc = 0.0f, gf = 0.0f;
        for (int pass = 0; pass * 8 < slots; ++pass) {
            R8 thd[kRowLayers], epd[kRowLayers], sgd[kRowLayers];
            for (int l = 0; l < lc; ++l) {
                thd[l] = sv(in.lt.thickness[lo + l], 2 + 3 * l, pass);
                epd[l] = sv(in.lt.eps[lo + l], 3 + 3 * l, pass);
                sgd[l] = sv(in.lt.sigma[lo + l], 4 + 3 * l, pass);
            }
            StackOut<R8> s;
            stack_row<R8>(sv(in.cos_theta[i], 0, pass), m, in.lt, thd, epd, sgd, sv((float)freq, 1, pass), s);
            float d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            auto add = [&](const float* gp, const R8& x) {
                if (!gp) return;
                const float w = gp[i];
                for (int j = 0; j < 8; ++j) d[j] += w * x.d[j];
            };
            add(g.r_te_re, s.r_te.re); add(g.r_te_im, s.r_te.im); add(g.r_tm_re, s.r_tm.re); add(g.r_tm_im, s.r_tm.im);
            add(g.t_te_re, s.t_te.re); add(g.t_te_im, s.t_te.im); add(g.t_tm_re, s.t_tm.re); add(g.t_tm_im, s.t_tm.im);
            add(g.R_te, s.R_te); add(g.R_tm, s.R_tm); add(g.T_te, s.T_te); add(g.T_tm, s.T_tm);
            for (int j = 0; j < 8; ++j) {
                const int sl = 8 * pass + j;
                if (sl >= slots) break;
                if (sl == 0) gc = d[j];
                else if (sl == 1) gf = d[j];
                else if (d[j] != 0.0f) {
                    const int l = (sl - 2) / 3, which = (sl - 2) % 3;
                    double* dst = which == 0 ? g_thick : which == 1 ? g_eps : g_sig;
                    if (dst) unsafeAtomicAdd(dst + lo + l, (double)d[j]);
                }
            }
        }
        if (g_cos) g_cos[i] = gc;
        g_freq_row[i] = gf;
    }
}

// ------------------------------------------------------------------ reductions
// sum of n floats in a fixed order (one block): strided float64 partials, tree
__global__ void __launch_bounds__(256) sum_rows_kernel(const float* __restrict__ x, int n, float* __restrict__ out) {
    __shared__ double part[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += (double)x[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)part[0];
}

__global__ void cast_layers_kernel(const double* __restrict__ acc, int n, float* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) out[i] = (float)acc[i];
}

__global__ void zero_doubles_kernel(double* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0.0;
}

int row_blocks(int n) { return std::max(1, std::min((n + 127) / 128, 65536)); }

FieldOutPtrs out_ptrs(const wc_field_rows_out* o) {
    return {o->field_vector, o->coefficient, o->path_field, o->path_gain, o->path_length_m, o->delay_s, o->direction};
}
FieldGrads grad_ptrs(const wc_field_rows_grad* g) {
    return {g->field_vector, g->coefficient, g->path_field, g->path_gain, g->path_length_m, g->delay_s,
            g->direction};
}

int check_out(const wc_field_rows_out* o, bool need_direction, const char* who) {
    if (!o || !o->field_vector || !o->coefficient || !o->path_field || !o->path_gain || !o->path_length_m ||
        !o->delay_s || (need_direction && !o->direction))
        return wc_fail("%s: output rows are required", who);
    return 0;
}

LayerTables layer_tables(const wc_layer_tables* t) {
    return {t->layer_offset, t->layer_count, t->layer_thickness, t->layer_eps, t->layer_sigma, t->layer_mu,
            t->material_count, t->layer_total};
}

int check_layers(const wc_layer_tables* t, const char* who) {
    if (!t || !t->layer_offset || !t->layer_count || (t->layer_total > 0 && (!t->layer_thickness || !t->layer_eps ||
                                                                            !t->layer_sigma || !t->layer_mu)))
        return wc_fail("%s: layer tables are required", who);
    return 0;
}

// per-layer float64 accumulators of the layer gradients (zeroed) and the
// row partials of the frequency gradient
struct ScratchView {
    double* thick;
    double* eps;
    double* sig;
    float* freq_row;
};
ScratchView scratch_view(void* ws, int rows, int layers) {
    char* p = static_cast<char*>(ws);
    ScratchView v;
    v.thick = reinterpret_cast<double*>(p);
    v.eps = v.thick + layers;
    v.sig = v.eps + layers;
    v.freq_row = reinterpret_cast<float*>(v.sig + layers);
    (void)rows;
    return v;
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" size_t wc_field_rows_workspace_bytes(int rows, int layers) {
    return (size_t)3 * 8 * (size_t)std::max(layers, 0) + 4 * (size_t)std::max(rows, 1) + 256;
}

// ------------------------------------------------------------------ free space
extern "C" int wc_field_free_space_jvp(int count, const float* source, const float* target, const float* tx_power,
                                       const float* tx_polarization, const float* rx_polarization,
                                       double frequency_hz, float tangent_frequency, const float* tangent_source,
                                       const float* tangent_target, const wc_field_rows_out* tangents,
                                       void* stream) {
    const char* who = "wc_field_free_space_jvp";
    if (count < 0) return wc_fail("%s: count must be non-negative", who);
    if (!(frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    if (int rc = check_out(tangents, true, who)) return rc;
    if (count == 0) return 0;
    hipLaunchKernelGGL(fs_jvp_kernel, dim3(row_blocks(count)), dim3(128), 0, (hipStream_t)stream, count, source,
                       target, tx_power, tx_polarization, rx_polarization, frequency_hz, tangent_frequency,
                       tangent_source, tangent_target, out_ptrs(tangents));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_field_free_space_vjp(int count, const float* source, const float* target, const float* tx_power,
                                       const float* tx_polarization, const float* rx_polarization,
                                       double frequency_hz, const wc_field_rows_grad* grads, float* grad_source,
                                       float* grad_target, float* grad_frequency, void* workspace, void* stream) {
    const char* who = "wc_field_free_space_vjp";
    if (count < 0) return wc_fail("%s: count must be non-negative", who);
    if (!(frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    if (!grads || !grad_frequency || !workspace) return wc_fail("%s: null argument", who);
    hipStream_t st = (hipStream_t)stream;
    const ScratchView v = scratch_view(workspace, count, 0);
    if (count > 0)
        hipLaunchKernelGGL(fs_vjp_kernel, dim3(row_blocks(count)), dim3(128), 0, st, count, source, target, tx_power,
                           tx_polarization, rx_polarization, frequency_hz, grad_ptrs(grads), grad_source, grad_target,
                           v.freq_row);
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(256), 0, st, v.freq_row, count, grad_frequency);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

// ------------------------------------------------------------------ reflection sequence
static ReflIn refl_in(const wc_reflection_rows* r) {
    return {r->depth, r->source, r->target, r->interaction_positions, r->interaction_normals, r->tx_power,
            r->tx_polarization, r->rx_polarization, r->eps_r, r->sigma_e, r->mu_r, r->gain, r->thickness};
}

static int check_refl(const wc_reflection_rows* r, const char* who) {
    if (!r || r->count < 0) return wc_fail("%s: invalid rows", who);
    if (r->depth < 1 || r->depth > kMaxDepth) return wc_fail("reflection depth must be in 1..%d", kMaxDepth);
    if (!(r->frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    if (r->count > 0 && (!r->source || !r->target || !r->interaction_positions || !r->interaction_normals ||
                         !r->tx_power || !r->tx_polarization || !r->rx_polarization || !r->eps_r || !r->sigma_e ||
                         !r->mu_r || !r->gain || !r->thickness))
        return wc_fail("%s: row inputs are required", who);
    return 0;
}

extern "C" int wc_field_reflection_sequence_jvp(const wc_reflection_rows* rows, const wc_reflection_rows* tangents,
                                                float tangent_frequency, const wc_field_rows_out* out, void* stream) {
    const char* who = "wc_field_reflection_sequence_jvp";
    if (int rc = check_refl(rows, who)) return rc;
    if (!tangents) return wc_fail("%s: null tangents", who);
    if (int rc = check_out(out, true, who)) return rc;
    if (rows->count == 0) return 0;
    ReflIn t = refl_in(tangents);
    t.depth = rows->depth;
    hipLaunchKernelGGL(refl_jvp_kernel, dim3(row_blocks(rows->count)), dim3(128), 0, (hipStream_t)stream,
                       rows->count, refl_in(rows), rows->frequency_hz, tangent_frequency, t, out_ptrs(out));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_field_reflection_sequence_vjp(const wc_reflection_rows* rows, const wc_field_rows_grad* grads,
                                                const wc_reflection_rows_gradout* out, void* workspace, void* stream) {
    const char* who = "wc_field_reflection_sequence_vjp";
    if (int rc = check_refl(rows, who)) return rc;
    if (!grads || !out || !out->grad_frequency || !workspace) return wc_fail("%s: null argument", who);
    hipStream_t st = (hipStream_t)stream;
    const ScratchView v = scratch_view(workspace, rows->count, 0);
    const ReflGrads g{out->grad_source, out->grad_target, out->grad_interaction_positions,
                      out->grad_interaction_normals, out->grad_eps_r, out->grad_sigma_e, out->grad_gain,
                      out->grad_thickness, v.freq_row};
    if (rows->count > 0)
        hipLaunchKernelGGL(refl_vjp_kernel, dim3(row_blocks(rows->count)), dim3(128), 0, st, rows->count,
                           refl_in(rows), rows->frequency_hz, grad_ptrs(grads), g);
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(256), 0, st, v.freq_row, rows->count, out->grad_frequency);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

// ------------------------------------------------------------------ transmission sequence
static int check_trans(const wc_transmission_rows* r, const char* who) {
    if (!r || r->count < 0) return wc_fail("%s: invalid rows", who);
    if (r->depth < 1 || r->depth > kMaxHits) return wc_fail("transmission depth must be in 1..%d", kMaxHits);
    if (!(r->frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    if (int rc = check_layers(&r->layers, who)) return rc;
    if (r->count > 0 && (!r->path_valid || !r->interaction_valid || !r->interaction_material_id || !r->source ||
                         !r->target || !r->interaction_normals || !r->tx_power || !r->tx_polarization ||
                         !r->rx_polarization))
        return wc_fail("%s: row inputs are required", who);
    return 0;
}

static TransIn trans_in(const wc_transmission_rows* r) {
    TransIn in;
    in.depth = r->depth;
    in.path_valid = r->path_valid;
    in.ivalid = r->interaction_valid;
    in.mat = r->interaction_material_id;
    in.src = r->source;
    in.tgt = r->target;
    in.nrm = r->interaction_normals;
    in.power = r->tx_power;
    in.txp = r->tx_polarization;
    in.rxp = r->rx_polarization;
    in.lt = layer_tables(&r->layers);
    return in;
}

extern "C" int wc_field_transmission_sequence(const wc_transmission_rows* rows, const wc_field_rows_out* out,
                                              void* stream) {
    const char* who = "wc_field_transmission_sequence";
    if (int rc = check_trans(rows, who)) return rc;
    if (int rc = check_out(out, true, who)) return rc;
    if (rows->count == 0) return 0;
    hipLaunchKernelGGL(trans_fwd_kernel, dim3(row_blocks(rows->count)), dim3(128), 0, (hipStream_t)stream,
                       rows->count, trans_in(rows), rows->frequency_hz, out_ptrs(out));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_field_transmission_sequence_jvp(const wc_transmission_rows* rows, const float* tangent_source,
                                                  const float* tangent_target, const float* tangent_normals,
                                                  const float* tangent_layer_thickness,
                                                  const float* tangent_layer_eps_r,
                                                  const float* tangent_layer_sigma_e, float tangent_frequency,
                                                  const wc_field_rows_out* out, void* stream) {
    const char* who = "wc_field_transmission_sequence_jvp";
    if (int rc = check_trans(rows, who)) return rc;
    if (int rc = check_out(out, false, who)) return rc;
    if (rows->count == 0) return 0;
    const TransTangents t{tangent_source, tangent_target, tangent_normals, tangent_layer_thickness,
                          tangent_layer_eps_r, tangent_layer_sigma_e};
    hipLaunchKernelGGL(trans_jvp_kernel, dim3(row_blocks(rows->count)), dim3(128), 0, (hipStream_t)stream,
                       rows->count, trans_in(rows), rows->frequency_hz, tangent_frequency, t, out_ptrs(out));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_field_transmission_sequence_vjp(const wc_transmission_rows* rows, const wc_field_rows_grad* grads,
                                                  float* grad_source, float* grad_target, float* grad_normals,
                                                  float* grad_layer_thickness, float* grad_layer_eps_r,
                                                  float* grad_layer_sigma_e, float* grad_frequency, void* workspace,
                                                  void* stream) {
    const char* who = "wc_field_transmission_sequence_vjp";
    if (int rc = check_trans(rows, who)) return rc;
    if (!grads || !grad_frequency || !workspace) return wc_fail("%s: null argument", who);
    hipStream_t st = (hipStream_t)stream;
    const int layers = rows->layers.layer_total;
    const ScratchView v = scratch_view(workspace, rows->count, layers);
    if (layers > 0)
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(std::max(1, (3 * layers + 255) / 256)), dim3(256), 0, st,
                           v.thick, 3 * layers);
    const TransGrads g{grad_source, grad_target, grad_normals, v.freq_row, v.thick, v.eps, v.sig};
    if (rows->count > 0)
        hipLaunchKernelGGL(trans_vjp_kernel, dim3(row_blocks(rows->count)), dim3(128), 0, st, rows->count,
                           trans_in(rows), rows->frequency_hz, grad_ptrs(grads), g);
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(256), 0, st, v.freq_row, rows->count, grad_frequency);
    if (layers > 0) {
        const int b = std::max(1, (layers + 255) / 256);
        if (grad_layer_thickness) hipLaunchKernelGGL(cast_layers_kernel, dim3(b), dim3(256), 0, st, v.thick, layers,
                                                     grad_layer_thickness);
        if (grad_layer_eps_r) hipLaunchKernelGGL(cast_layers_kernel, dim3(b), dim3(256), 0, st, v.eps, layers,
                                                 grad_layer_eps_r);
        if (grad_layer_sigma_e) hipLaunchKernelGGL(cast_layers_kernel, dim3(b), dim3(256), 0, st, v.sig, layers,
                                                   grad_layer_sigma_e);
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

// ------------------------------------------------------------------ layer stack
static int check_stack(int count, const float* cos_theta, const int* mat, const wc_layer_tables* lt, double freq,
                       const char* who) {
    if (count < 0) return wc_fail("%s: count must be non-negative", who);
    if (!(freq > 0.0)) return wc_fail("frequency_hz must be positive");
    if (int rc = check_layers(lt, who)) return rc;
    if (count > 0 && (!cos_theta || !mat)) return wc_fail("%s: cos_theta and material_id are required", who);
    return 0;
}

static StackPtrs stack_ptrs(const wc_layer_stack_out* o) {
    return {o->r_te_real, o->r_te_imag, o->r_tm_real, o->r_tm_imag, o->t_te_real, o->t_te_imag, o->t_tm_real,
            o->t_tm_imag, o->cap_R_te, o->cap_R_tm, o->cap_T_te, o->cap_T_tm};
}

static bool stack_out_ok(const wc_layer_stack_out* o) {
    return o && o->r_te_real && o->r_te_imag && o->r_tm_real && o->r_tm_imag && o->t_te_real && o->t_te_imag &&
           o->t_tm_real && o->t_tm_imag && o->cap_R_te && o->cap_R_tm && o->cap_T_te && o->cap_T_tm;
}

extern "C" int wc_em_layer_stack_eval(int count, const float* cos_theta, const int* material_id,
                                      const wc_layer_tables* layers, double frequency_hz,
                                      const wc_layer_stack_out* out, void* stream) {
    const char* who = "wc_em_layer_stack_eval";
    if (int rc = check_stack(count, cos_theta, material_id, layers, frequency_hz, who)) return rc;
    if (!stack_out_ok(out)) return wc_fail("%s: every output is required", who);
    if (count == 0) return 0;
    const StackIn in{cos_theta, material_id, layer_tables(layers)};
    hipLaunchKernelGGL(stack_fwd_kernel, dim3(row_blocks(count)), dim3(128), 0, (hipStream_t)stream, count, in,
                       frequency_hz, stack_ptrs(out));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_em_layer_stack_jvp(int count, const float* cos_theta, const int* material_id,
                                     const wc_layer_tables* layers, double frequency_hz,
                                     const float* tangent_cos_theta, const float* tangent_layer_thickness,
                                     const float* tangent_layer_eps_r, const float* tangent_layer_sigma_e,
                                     float tangent_frequency, const wc_layer_stack_out* out, void* stream) {
    const char* who = "wc_em_layer_stack_jvp";
    if (int rc = check_stack(count, cos_theta, material_id, layers, frequency_hz, who)) return rc;
    if (!stack_out_ok(out)) return wc_fail("%s: every output is required", who);
    if (count == 0) return 0;
    const StackIn in{cos_theta, material_id, layer_tables(layers)};
    const StackTangents t{tangent_cos_theta, tangent_layer_thickness, tangent_layer_eps_r, tangent_layer_sigma_e};
    hipLaunchKernelGGL(stack_jvp_kernel, dim3(row_blocks(count)), dim3(128), 0, (hipStream_t)stream, count, in,
                       frequency_hz, tangent_frequency, t, stack_ptrs(out));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_em_layer_stack_vjp(int count, const float* cos_theta, const int* material_id,
                                     const wc_layer_tables* layers, double frequency_hz,
                                     const wc_layer_stack_out* grads, float* grad_cos_theta,
                                     float* grad_layer_thickness, float* grad_layer_eps_r,
                                     float* grad_layer_sigma_e, float* grad_frequency, void* workspace,
                                     void* stream) {
    const char* who = "wc_em_layer_stack_vjp";
    if (int rc = check_stack(count, cos_theta, material_id, layers, frequency_hz, who)) return rc;
    if (!grads || !grad_frequency || !workspace) return wc_fail("%s: null argument", who);
    hipStream_t st = (hipStream_t)stream;
    const int nl = layers->layer_total;
    const ScratchView v = scratch_view(workspace, count, nl);
    if (nl > 0)
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(std::max(1, (3 * nl + 255) / 256)), dim3(256), 0, st, v.thick,
                           3 * nl);
    const StackIn in{cos_theta, material_id, layer_tables(layers)};
    const StackGradIn g{grads->r_te_real, grads->r_te_imag, grads->r_tm_real, grads->r_tm_imag, grads->t_te_real,
                        grads->t_te_imag, grads->t_tm_real, grads->t_tm_imag, grads->cap_R_te, grads->cap_R_tm,
                        grads->cap_T_te, grads->cap_T_tm};
    if (count > 0)
        hipLaunchKernelGGL(stack_vjp_kernel, dim3(row_blocks(count)), dim3(128), 0, st, count, in, frequency_hz, g,
                           grad_cos_theta, v.freq_row, v.thick, v.eps, v.sig);
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(256), 0, st, v.freq_row, count, grad_frequency);
    if (nl > 0) {
        const int b = std::max(1, (nl + 255) / 256);
        if (grad_layer_thickness)
            hipLaunchKernelGGL(cast_layers_kernel, dim3(b), dim3(256), 0, st, v.thick, nl, grad_layer_thickness);
        if (grad_layer_eps_r) hipLaunchKernelGGL(cast_layers_kernel, dim3(b), dim3(256), 0, st, v.eps, nl, grad_layer_eps_r);
        if (grad_layer_sigma_e)
            hipLaunchKernelGGL(cast_layers_kernel, dim3(b), dim3(256), 0, st, v.sig, nl, grad_layer_sigma_e);
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}
