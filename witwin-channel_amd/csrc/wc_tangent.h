// Tangent (forward-mode) path evaluators used by the fixed-topology adjoint.
// Each evaluator recomputes one stored path — interaction geometry included —
// from the transmitter position, the frequency and the per-interaction
// material parameters carried as dual numbers, so d(field)/d(parameter) is
// exact for the frozen topology. Values follow the forward kernels of
// wc_paths.h; only the tangents are consumed.
#pragma once

#include "wc_paths.h"

namespace wc {

template <typename R> __device__ __forceinline__ V3<R> lift3(const V3<float>& v) {
    return {lift<R>(v.x), lift<R>(v.y), lift<R>(v.z)};
}

template <typename R> __device__ __forceinline__ V3<R> scale3(const V3<float>& v, const R& s) {
    return {s * v.x, s * v.y, s * v.z};
}

// ----------------------------------------------------------------- line of sight
template <typename R>
__device__ inline C3<R> los_tangent(const V3<R>& tx, const V3<float>& rx, const V3<float>& tx_pol, const R& freq,
                                    double k_d, double dist_d, V3<R>& arrival) {
    const V3<R> seg = lift3<R>(rx) - tx;
    const R dist = norm(seg);
    const V3<R> dir = seg * (lift<R>(1.0f) / dist);
    const R k = wave_number(freq);
    const R amp = lift<R>(1.0f) / (k * dist * 2.0f);
    const R ph = PhaseOf<R>::make(k, dist, k_d, dist_d);
    arrival = dir;
    return c3_axis(transverse(dir, lift3<R>(tx_pol)), scale(cexpj(ph), amp));
}

// ----------------------------------------------------------------- reflection
template <typename R>
__device__ inline C3<R> reflection_tangent(const SceneView& s, int depth, const int* groups, const int* faces,
                                           const V3<R>& tx, const V3<float>& rx, const V3<float>& tx_pol,
                                           const R& freq, const R* eps, const R* sig, double k_d, double length_d,
                                           V3<R>& arrival) {
    V3<R> img[kMaxDepth];
    V3<float> nrm[kMaxDepth];
    float off[kMaxDepth];
    V3<R> p = tx;
    for (int j = 0; j < depth; ++j) {
        double o;
        nrm[j] = to_f(group_normal(s, groups[j], o));
        off[j] = (float)o;
        const R dist = dot(p, lift3<R>(nrm[j])) - off[j];
        p = p - scale3<R>(nrm[j], dist * 2.0f);
        img[j] = p;
    }
    V3<R> hits[kMaxDepth];
    V3<R> next = lift3<R>(rx);
    for (int j = depth - 1; j >= 0; --j) {
        const V3<R> n = lift3<R>(nrm[j]);
        const R si = dot(n, img[j]) - off[j];
        const R sn = dot(n, next) - off[j];
        const R t = si / (si - sn);
        hits[j] = img[j] + (next - img[j]) * t;
        next = hits[j];
    }
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    V3<R> prev = tx;
    V3<R> incident = safe_normalize(hits[0] - prev, ez);
    C3<R> value = c3_from_real(transverse(incident, lift3<R>(tx_pol)));
    V3<R> outgoing = incident;
    for (int b = 0; b < depth; ++b) {
        incident = safe_normalize(hits[b] - prev, outgoing);
        const int m = s.face_material[faces[b]];
        value = reflect_c3(value, incident, lift3<R>(nrm[b]), eps[b], sig[b], lift<R>(s.mat_mu[m]),
                           lift<R>(s.mat_gain[m]), lift<R>(s.mat_thickness[m]), freq, outgoing);
        prev = hits[b];
    }
    arrival = safe_normalize(lift3<R>(rx) - prev, outgoing);
    const R L = norm(lift3<R>(rx) - img[depth - 1]);
    const R k = wave_number(freq);
    const R amp = lift<R>(1.0f) / (k * L * 2.0f);
    return c3_scale(value, scale(cexpj(PhaseOf<R>::make(k, L, k_d, length_d)), amp));
}

// ----------------------------------------------------------------- transmission
template <typename R>
__device__ inline C3<R> transmission_tangent(const SceneView& s, int count, const Hit* hits, const V3<R>& tx,
                                             const V3<float>& rx, const V3<float>& tx_pol, const R& freq,
                                             const R* eps, const R* sig, double k_d, double carrier_d,
                                             V3<R>& arrival) {
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    const V3<R> seg = lift3<R>(rx) - tx;
    const R dist = norm(seg);
    const V3<R> dir = seg * (lift<R>(1.0f) / dist);
    C3<R> value = c3_from_real(transverse(dir, lift3<R>(tx_pol)));
    R carrier = dist;
    for (int w = 0; w < count; ++w) {
        double o;
        V3<R> n = safe_normalize(lift3<R>(to_f(group_normal(s, hits[w].group, o))), ez);
        if (val(dot(dir, n)) > 0.0f) n = -n;
        const R ct = w_min(w_max(w_abs(dot(dir, n)), kSmallEpsF), 1.0f);
        const int m = s.face_material[hits[w].face];
        const int lo = s.mat_layer_offset[m], lc = s.mat_layer_count[m];
        // the material parameters of the first layer carry the tangents
        R eps_layers[4], sig_layers[4];
        const int lcc = lc < 4 ? lc : 4;
        for (int l = 0; l < lcc; ++l) {
            eps_layers[l] = l == 0 ? eps[w] : lift<R>(s.layer_eps[lo + l]);
            sig_layers[l] = l == 0 ? sig[w] : lift<R>(s.layer_sigma[lo + l]);
        }
        const Cx<R> t_te = stack_t<R>(ct, s.layer_thickness + lo, s.layer_eps + lo, s.layer_sigma + lo,
                                      s.layer_mu + lo, lcc, freq, 0, eps_layers, sig_layers);
        const Cx<R> t_tm = stack_t<R>(ct, s.layer_thickness + lo, s.layer_eps + lo, s.layer_sigma + lo,
                                      s.layer_mu + lo, lcc, freq, 1, eps_layers, sig_layers);
        const V3<R> sax = safe_normalize(cross(n, dir), stable_perp_basis(dir, n));
        const V3<R> pax = safe_normalize(cross(sax, dir), stable_perp_basis(dir, sax));
        const Cx<R> e_s = c3_dot_real(value, sax);
        const Cx<R> e_p = c3_dot_real(value, pax);
        value = c3_axis(sax, t_te * e_s) + c3_axis(pax, t_tm * e_p);
        float wall = 0.0f;
        for (int l = 0; l < lc; ++l) wall += fmaxf(s.layer_thickness[lo + l], 0.0f);
        carrier = carrier - ct * wall;
    }
    arrival = dir;
    const R k = wave_number(freq);
    const R amp = lift<R>(1.0f) / (k * dist * 2.0f);
    return c3_scale(value, scale(cexpj(PhaseOf<R>::make(k, carrier, k_d, carrier_d)), amp));
}

// ----------------------------------------------------------------- diffraction
// value of one cot*F term in double plus its partial derivatives in beta and
// kL (and, when dn_re is given, in the wedge index n)
__device__ inline void utd_term_partials(double n, double beta, int sgn, double kL, double& re, double& im,
                                         double& db_re, double& db_im, double& dk_re, double& dk_im,
                                         double* dn_re = nullptr, double* dn_im = nullptr) {
    const double arg = (kPi + sgn * beta) / (2.0 * n);
    double sa, ca;
    sincos(arg, &sa, &ca);
    if (fabs(sa) < 1.0e-9) {
        // shadow-boundary limit n sqrt(2 pi kL) sgn(eps) e^{j pi/4}: d/dkL only
        utd_term(n, beta, sgn, kL, re, im);
        const double f = 0.5 / fmax(kL, 1e-30);
        dk_re = re * f;
        dk_im = im * f;
        db_re = 0.0;
        db_im = 0.0;
        if (dn_re) {
            // the limit is proportional to n
            *dn_re = re / n;
            *dn_im = im / n;
        }
        return;
    }
    const double N = nearbyint((beta + sgn * kPi) / (2.0 * n * kPi));
    // a = 2 cos^2(h), h = (2 pi n N - beta) / 2; da/dbeta = sin(2h) = 2 sin(h) cos(h)
    double sh, c;
    sincos((2.0 * n * kPi * N - beta) * 0.5, &sh, &c);
    const double a = 2.0 * c * c;
    const double da = 2.0 * sh * c;
    const double X = kL * a;
    const double cot = ca / sa;
    const double dcot = -(sgn / (2.0 * n)) / (sa * sa);
    double f_re, f_im;
    utd_transition(X, f_re, f_im);
    // the value of utd_term, from the one transition-function evaluation
    re = cot * f_re;
    im = cot * f_im;
    // F'(X) = (j + 1/(2X)) F - j
    double fp_re = 0.0, fp_im = 0.0;
    if (X > 0.0) {
        const double h = 0.5 / X;
        fp_re = h * f_re - f_im;
        fp_im = h * f_im + f_re - 1.0;
    }
    db_re = dcot * f_re + cot * fp_re * kL * da;
    db_im = dcot * f_im + cot * fp_im * kL * da;
    dk_re = cot * fp_re * a;
    dk_im = cot * fp_im * a;
    if (dn_re) {
        // d arg / dn = -arg / n; da/dn = -4 pi N cos(h) sin(h)
        const double dcot_n = arg / (n * sa * sa);
        const double da_n = -4.0 * kPi * N * c * sh;
        *dn_re = dcot_n * f_re + cot * fp_re * kL * da_n;
        *dn_im = dcot_n * f_im + cot * fp_im * kL * da_n;
    }
}

template <typename R> struct UtdTangent;
template <int N> struct UtdTangent<Dual<N>> {
    using D = Dual<N>;
    __device__ static Cx<D> term(double n, const D& beta, int sgn, const D& kL) {
        return term_at(n, (double)beta.v, beta, sgn, kL);
    }
    // the term at the double argument beta_v with the tangents of beta_t
    __device__ static Cx<D> term_at(double n, double beta_v, const D& beta_t, int sgn, const D& kL) {
        double re, im, br, bi, kr, ki;
        utd_term_partials(n, beta_v, sgn, (double)kL.v, re, im, br, bi, kr, ki);
        Cx<D> out;
        out.re.v = (float)re;
        out.im.v = (float)im;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            out.re.d[i] = (float)(br * beta_t.d[i] + kr * kL.d[i]);
            out.im.d[i] = (float)(bi * beta_t.d[i] + ki * kL.d[i]);
        }
        return out;
    }
    // the same term with the wedge index n carrying tangents too
    __device__ static Cx<D> term_n(const D& n, const D& beta, int sgn, const D& kL) {
        return term_n_at(n, (double)beta.v, beta, sgn, kL);
    }
    __device__ static Cx<D> term_n_at(const D& n, double beta_v, const D& beta_t, int sgn, const D& kL) {
        double re, im, br, bi, kr, ki, nr, ni;
        utd_term_partials((double)n.v, beta_v, sgn, (double)kL.v, re, im, br, bi, kr, ki, &nr, &ni);
        Cx<D> out;
        out.re.v = (float)re;
        out.im.v = (float)im;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            out.re.d[i] = (float)(br * beta_t.d[i] + kr * kL.d[i] + nr * n.d[i]);
            out.im.d[i] = (float)(bi * beta_t.d[i] + ki * kL.d[i] + ni * n.d[i]);
        }
        return out;
    }
    // E1d assembly of one term t (utd_term_e1d): t T for an inactive term;
    // for the active one (t - odd_w) T + odd_w gamma with the mirrored term
    // evaluated at 2 b* - beta in double. which: 0 -> b* = -pi, 1 -> +pi,
    // 2 -> pi, 3 -> (2n - 1) pi, whose tangent 2 pi dn is carried when
    // n_tangent is set.
    __device__ static Cx<D> mend(const Cx<D>& t, bool active, const D& n, bool n_tangent, const D& beta, int sgn,
                                 int which, const D& kL, const Cx<D>& trunc, const D& gamma) {
        if (!active) return t * trunc;
        const double nv = (double)n.v;
        const double bstar = which == 0 ? -kPi : (which == 3 ? (2.0 * nv - 1.0) * kPi : kPi);
        D delta, mirror_t;
        delta.v = (float)((double)beta.v - bstar);
        mirror_t.v = 0.0f;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const float db = (which == 3 && n_tangent) ? (float)kTwoPi * n.d[i] : 0.0f;
            delta.d[i] = beta.d[i] - db;
            mirror_t.d[i] = 2.0f * db - beta.d[i];
        }
        const D dw = w_sqrt(make_dual<N>((float)kTwoPi) / kL) * 4.0f;
        const D x = delta / dw;
        const D w = w_exp(-(x * x));
        if (!(w.v > 1.0e-20f)) return t * trunc;
        const double mirror_v = 2.0 * bstar - (double)beta.v;
        const Cx<D> tm = n_tangent ? term_n_at(n, mirror_v, mirror_t, sgn, kL)
                                   : term_at(nv, mirror_v, mirror_t, sgn, kL);
        const Cx<D> odd = scale(t - tm, w * 0.5f);
        return (t - odd) * trunc + scale(odd, gamma);
    }
    // corner blend gamma = 1 / (1 + exp(u / 0.1)), u = sg max(-t*, t* - l)
    __device__ static D corner_blend(const D& sg, const D& tstar, const D& length) {
        const D a = -tstar, b = tstar - length;
        const D u = sg * (a.v >= b.v ? a : b);
        if (u.v * 10.0f > 80.0f) return make_dual<N>(0.0f);
        return make_dual<N>(1.0f) / (w_exp(u * 10.0f) + 1.0f);
    }
    // Fr(x) = C(x) + jS(x), dFr/dx = e^{j pi x^2 / 2}
    __device__ static Cx<D> fresnel(const D& x) {
        double c, s, dr, di;
        fresnel_complex((double)x.v, c, s, &dr, &di);
        const float gr = (float)dr, gi = (float)di;
        Cx<D> out;
        out.re.v = (float)c;
        out.im.v = (float)s;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            out.re.d[i] = gr * x.d[i];
            out.im.d[i] = gi * x.d[i];
        }
        return out;
    }
};

template <typename R>
__device__ inline R wedge_angle_t(const V3<R>& u, const V3<float>& t0, const V3<float>& n0) {
    R a = w_atan2(dot(u, lift3<R>(n0)), dot(u, lift3<R>(t0)));
    if (val(a) < 0.0f) a = a + (float)kTwoPi;
    return a;
}

template <typename R> __device__ __forceinline__ Cx<float> cval(const Cx<R>& a) { return {val(a.re), val(a.im)}; }
template <typename R> __device__ __forceinline__ V3<float> vval(const V3<R>& a) {
    return {val(a.x), val(a.y), val(a.z)};
}

// The diffracted field is linear in the four face reflection coefficients
// (r0_s, r0_h on face 0, rn_s, rn_h on face n), the only place the face
// materials enter. Their partials (values only) and the incidence cosines
// handed to each face's slab coefficients let the adjoint differentiate the
// materials without carrying material tangents through the whole path.
struct DiffFaceTerms {
    C3<float> d_r[4];  // d field / d r0_s, r0_h, rn_s, rn_h
    float cos_face[2];
};

template <typename R>
__device__ inline C3<R> diffraction_tangent(const SceneView& s, const Edge& e, const V3<R>& tx, const V3<float>& rx,
                                            const V3<float>& tx_pol, const R& freq, const R* eps, const R* sig,
                                            double k_d, double length_d, V3<R>& arrival,
                                            DiffFaceTerms* faces = nullptr) {
    const V3<float> A{e.ax, e.ay, e.az}, ed{e.dx, e.dy, e.dz};
    const V3<float> t0{e.t0x, e.t0y, e.t0z}, n0{e.n0x, e.n0y, e.n0z};
    const V3<R> AR = lift3<R>(A), EdR = lift3<R>(ed), RxR = lift3<R>(rx);
    const R a = dot(tx - AR, EdR);
    const R b = dot(RxR - AR, EdR);
    const V3<R> pt = tx - AR - EdR * a;
    const V3<R> pr = RxR - AR - EdR * b;
    const R rho_p = norm(pt), rho = norm(pr);
    const R tstar = a + (b - a) * rho_p / (rho_p + rho);
    const V3<R> q = AR + EdR * tstar;
    const R sp = norm(q - tx), so = norm(RxR - q);
    const V3<R> si = (q - tx) * (lift<R>(1.0f) / sp);
    const V3<R> sov = (RxR - q) * (lift<R>(1.0f) / so);
    const float npi = e.n * (float)kPi;
    const R phi_p = wedge_angle_t(pt, t0, n0);
    R phi = wedge_angle_t(pr, t0, n0);
    if (val(phi) > npi) phi = lift<R>((val(phi) - npi < (float)kTwoPi - val(phi)) ? npi : 0.0f);
    const R cb = dot(si, EdR);
    const R sin_b0 = w_sqrt(w_max(lift<R>(1.0f) - cb * cb, 1.0e-12f));
    V3<R> phi_i = -cross(EdR, si);
    phi_i = phi_i * (lift<R>(1.0f) / norm(phi_i));
    const V3<R> b0_i = cross(phi_i, si);
    V3<R> phi_o = cross(EdR, sov);
    phi_o = phi_o * (lift<R>(1.0f) / norm(phi_o));
    const V3<R> b0_o = cross(phi_o, sov);
    const V3<R> tpol = transverse(si, lift3<R>(tx_pol));
    const R e_b = dot(tpol, b0_i), e_p = dot(tpol, phi_i);
    const R L = so * sp * sin_b0 * sin_b0 / (so + sp);
    const R k = wave_number(freq);
    const int m0 = e.material0;
    const int mn = e.material1 >= 0 ? e.material1 : e.material0;
    Cx<R> r0_s, r0_h, rn_s, rn_h;
    const R cos0 = w_abs(w_sin(phi_p));
    const R cosn = w_abs(w_sin(lift<R>(npi) - phi));
    slab_fresnel<R>(cos0, eps[0], sig[0], lift<R>(s.mat_mu[m0]), lift<R>(s.mat_gain[m0]),
                    lift<R>(s.mat_thickness[m0]), freq, r0_s, r0_h);
    slab_fresnel<R>(cosn, eps[1], sig[1], lift<R>(s.mat_mu[mn]), lift<R>(s.mat_gain[mn]),
                    lift<R>(s.mat_thickness[mn]), freq, rn_s, rn_h);
    const double nn = (double)e.n;
    const R kL = k * L;
    // finite-edge truncation factor and corner blend (E1d, see utd_term_e1d)
    const R kappa = rho_p * rho_p / (sp * sp * sp) + rho * rho / (so * so * so);
    const R sg = w_sqrt(k * kappa * (float)(1.0 / kPi));
    const Cx<R> f1 = UtdTangent<R>::fresnel(sg * (lift<R>(e.length) - tstar));
    const Cx<R> f0 = UtdTangent<R>::fresnel(sg * (-tstar));
    const Cx<R> dfr = f1 - f0;
    const Cx<R> dconj{dfr.re, -dfr.im};
    const Cx<R> trunc = cxc<R>(0.5f, 0.5f) * dconj;
    const R gamma = UtdTangent<R>::corner_blend(sg, tstar, lift<R>(e.length));
    const UtdBoundaries bd = utd_boundaries(nn, (double)val(phi), (double)val(phi_p));
    const R n_r = lift<R>(e.n);
    const Cx<R> t1 = UtdTangent<R>::mend(UtdTangent<R>::term(nn, phi - phi_p, +1, kL), !bd.incident_up, n_r, false,
                                         phi - phi_p, +1, 0, kL, trunc, gamma);
    const Cx<R> t2 = UtdTangent<R>::mend(UtdTangent<R>::term(nn, phi - phi_p, -1, kL), bd.incident_up, n_r, false,
                                         phi - phi_p, -1, 1, kL, trunc, gamma);
    const Cx<R> t3 = UtdTangent<R>::mend(UtdTangent<R>::term(nn, phi + phi_p, -1, kL), bd.reflect_0, n_r, false,
                                         phi + phi_p, -1, 2, kL, trunc, gamma);
    const Cx<R> t4 = UtdTangent<R>::mend(UtdTangent<R>::term(nn, phi + phi_p, +1, kL), !bd.reflect_0, n_r, false,
                                         phi + phi_p, +1, 3, kL, trunc, gamma);
    // prefactor -e^{-j pi/4} / (2 n sqrt(2 pi k) sin b0)
    const R pre_mag = lift<R>(1.0f) / (w_sqrt(k * (float)kTwoPi) * sin_b0 * (2.0f * e.n));
    const Cx<R> pre = scale(Cx<R>{lift<R>(-(float)cos(-0.25 * kPi)), lift<R>(-(float)sin(-0.25 * kPi))}, pre_mag);
    const Cx<R> ds = pre * (t1 + t2 + r0_s * t3 + rn_s * t4);
    const Cx<R> dh = pre * (t1 + t2 + r0_h * t3 + rn_h * t4);
    const R amp = w_sqrt(sp / (so * (so + sp))) / (k * sp * 2.0f);
    const R ph = PhaseOf<R>::make(k, so + sp, k_d, length_d);
    const Cx<R> prop = scale(cexpj(ph), amp);
    arrival = sov;
    if (faces) {
        const Cx<float> pp = cval(pre) * cval(prop);
        const Cx<float> a3 = pp * cval(t3), a4 = pp * cval(t4);
        const V3<float> bo = vval(b0_o), po = vval(phi_o);
        const Cx<float> sb{-val(e_b), 0.0f}, sp_{-val(e_p), 0.0f};
        faces->d_r[0] = c3_axis(bo, a3 * sb);
        faces->d_r[1] = c3_axis(po, a3 * sp_);
        faces->d_r[2] = c3_axis(bo, a4 * sb);
        faces->d_r[3] = c3_axis(po, a4 * sp_);
        faces->cos_face[0] = val(cos0);
        faces->cos_face[1] = val(cosn);
    }
    return c3_axis(b0_o, (ds * prop) * Cx<R>{-e_b, lift<R>(0.f)}) + c3_axis(phi_o, (dh * prop) * Cx<R>{-e_p, lift<R>(0.f)});
}


// ----------------------------------------------------------------- geometry tangents
// Path evaluators whose interaction geometry carries the tangents (the vertex
// adjoint): the plane (n, c) of every reflecting group, the normal of every
// transmitted wall, and the origin, frame rotation, length and wedge index of
// a diffracting edge. Endpoints, frequency and materials are constants. The
// values equal the evaluators above; only the tangents are consumed.

template <typename R>
__device__ inline C3<R> reflection_tangent_planes(const SceneView& s, int depth, const int* faces, const V3<R>* nrm,
                                                  const R* off, const V3<float>& tx, const V3<float>& rx,
                                                  const V3<float>& tx_pol, float freq, double k_d, double length_d,
                                                  V3<R>& arrival) {
    V3<R> img[kMaxDepth];
    V3<R> p = lift3<R>(tx);
    for (int j = 0; j < depth; ++j) {
        const R dist = dot(p, nrm[j]) - off[j];
        p = p - nrm[j] * (dist * 2.0f);
        img[j] = p;
    }
    V3<R> hits[kMaxDepth];
    V3<R> next = lift3<R>(rx);
    for (int j = depth - 1; j >= 0; --j) {
        const R si = dot(nrm[j], img[j]) - off[j];
        const R sn = dot(nrm[j], next) - off[j];
        const R t = si / (si - sn);
        hits[j] = img[j] + (next - img[j]) * t;
        next = hits[j];
    }
    const R fr = lift<R>(freq);
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    V3<R> prev = lift3<R>(tx);
    V3<R> incident = safe_normalize(hits[0] - prev, ez);
    C3<R> value = c3_from_real(transverse(incident, lift3<R>(tx_pol)));
    V3<R> outgoing = incident;
    for (int b = 0; b < depth; ++b) {
        incident = safe_normalize(hits[b] - prev, outgoing);
        const int m = s.face_material[faces[b]];
        value = reflect_c3(value, incident, nrm[b], lift<R>(s.mat_eps[m]), lift<R>(s.mat_sigma[m]),
                           lift<R>(s.mat_mu[m]), lift<R>(s.mat_gain[m]), lift<R>(s.mat_thickness[m]), fr, outgoing);
        prev = hits[b];
    }
    arrival = safe_normalize(lift3<R>(rx) - prev, outgoing);
    const R L = norm(lift3<R>(rx) - img[depth - 1]);
    const R k = wave_number(fr);
    const R amp = lift<R>(1.0f) / (k * L * 2.0f);
    return c3_scale(value, scale(cexpj(PhaseOf<R>::make(k, L, k_d, length_d)), amp));
}

template <typename R>
__device__ inline C3<R> transmission_tangent_walls(const SceneView& s, int count, const Hit* hits,
                                                   const V3<R>* wall_n, const V3<float>& tx, const V3<float>& rx,
                                                   const V3<float>& tx_pol, float freq, double k_d, double carrier_d,
                                                   V3<R>& arrival) {
    const V3<R> ez = v3c<R>(0.f, 0.f, 1.f);
    const V3<R> seg = lift3<R>(rx - tx);
    const R dist = norm(seg);
    const V3<R> dir = seg * (lift<R>(1.0f) / dist);
    const R fr = lift<R>(freq);
    C3<R> value = c3_from_real(transverse(dir, lift3<R>(tx_pol)));
    R carrier = dist;
    for (int w = 0; w < count; ++w) {
        V3<R> n = safe_normalize(wall_n[w], ez);
        if (val(dot(dir, n)) > 0.0f) n = -n;
        const R ct = w_min(w_max(w_abs(dot(dir, n)), kSmallEpsF), 1.0f);
        const int m = s.face_material[hits[w].face];
        const int lo = s.mat_layer_offset[m], lc = s.mat_layer_count[m];
        const int lcc = lc < 4 ? lc : 4;
        const Cx<R> t_te = stack_t<R>(ct, s.layer_thickness + lo, s.layer_eps + lo, s.layer_sigma + lo,
                                      s.layer_mu + lo, lcc, fr, 0, nullptr, nullptr);
        const Cx<R> t_tm = stack_t<R>(ct, s.layer_thickness + lo, s.layer_eps + lo, s.layer_sigma + lo,
                                      s.layer_mu + lo, lcc, fr, 1, nullptr, nullptr);
        const V3<R> sax = safe_normalize(cross(n, dir), stable_perp_basis(dir, n));
        const V3<R> pax = safe_normalize(cross(sax, dir), stable_perp_basis(dir, sax));
        const Cx<R> e_s = c3_dot_real(value, sax);
        const Cx<R> e_p = c3_dot_real(value, pax);
        value = c3_axis(sax, t_te * e_s) + c3_axis(pax, t_tm * e_p);
        float wall = 0.0f;
        for (int l = 0; l < lc; ++l) wall += fmaxf(s.layer_thickness[lo + l], 0.0f);
        carrier = carrier - ct * wall;
    }
    arrival = dir;
    const R k = wave_number(fr);
    const R amp = lift<R>(1.0f) / (k * dist * 2.0f);
    return c3_scale(value, scale(cexpj(PhaseOf<R>::make(k, carrier, k_d, carrier_d)), amp));
}

// Edge geometry with tangents: origin A, unit direction d, face-0 frame
// (t0, n0), length and exterior wedge index n.
template <typename R> struct EdgeGeom {
    V3<R> A, d, t0, n0;
    R length, n;
};

template <typename R>
__device__ inline R wedge_angle_r(const V3<R>& u, const V3<R>& t0, const V3<R>& n0) {
    R a = w_atan2(dot(u, n0), dot(u, t0));
    if (val(a) < 0.0f) a = a + (float)kTwoPi;
    return a;
}

template <typename R>
__device__ inline C3<R> diffraction_tangent_edge(const SceneView& s, const Edge& e, const EdgeGeom<R>& g,
                                                 const V3<float>& tx, const V3<float>& rx, const V3<float>& tx_pol,
                                                 float freq, double k_d, double length_d, V3<R>& arrival) {
    const V3<R> TxR = lift3<R>(tx), RxR = lift3<R>(rx);
    const R fr = lift<R>(freq);
    const R a = dot(TxR - g.A, g.d);
    const R b = dot(RxR - g.A, g.d);
    const V3<R> pt = TxR - g.A - g.d * a;
    const V3<R> pr = RxR - g.A - g.d * b;
    const R rho_p = norm(pt), rho = norm(pr);
    const R tstar = a + (b - a) * rho_p / (rho_p + rho);
    const V3<R> q = g.A + g.d * tstar;
    const R sp = norm(q - TxR), so = norm(RxR - q);
    const V3<R> si = (q - TxR) * (lift<R>(1.0f) / sp);
    const V3<R> sov = (RxR - q) * (lift<R>(1.0f) / so);
    const R npi = g.n * (float)kPi;
    const R phi_p = wedge_angle_r(pt, g.t0, g.n0);
    R phi = wedge_angle_r(pr, g.t0, g.n0);
    if (val(phi) > val(npi)) phi = (val(phi) - val(npi) < (float)kTwoPi - val(phi)) ? npi : lift<R>(0.0f);
    const R cb = dot(si, g.d);
    const R sin_b0 = w_sqrt(w_max(lift<R>(1.0f) - cb * cb, 1.0e-12f));
    V3<R> phi_i = -cross(g.d, si);
    phi_i = phi_i * (lift<R>(1.0f) / norm(phi_i));
    const V3<R> b0_i = cross(phi_i, si);
    V3<R> phi_o = cross(g.d, sov);
    phi_o = phi_o * (lift<R>(1.0f) / norm(phi_o));
    const V3<R> b0_o = cross(phi_o, sov);
    const V3<R> tpol = transverse(si, lift3<R>(tx_pol));
    const R e_b = dot(tpol, b0_i), e_p = dot(tpol, phi_i);
    const R L = so * sp * sin_b0 * sin_b0 / (so + sp);
    const R k = wave_number(fr);
    const int m0 = e.material0;
    const int mn = e.material1 >= 0 ? e.material1 : e.material0;
    Cx<R> r0_s, r0_h, rn_s, rn_h;
    const R cos0 = w_abs(w_sin(phi_p));
    const R cosn = w_abs(w_sin(npi - phi));
    slab_fresnel<R>(cos0, lift<R>(s.mat_eps[m0]), lift<R>(s.mat_sigma[m0]), lift<R>(s.mat_mu[m0]),
                    lift<R>(s.mat_gain[m0]), lift<R>(s.mat_thickness[m0]), fr, r0_s, r0_h);
    slab_fresnel<R>(cosn, lift<R>(s.mat_eps[mn]), lift<R>(s.mat_sigma[mn]), lift<R>(s.mat_mu[mn]),
                    lift<R>(s.mat_gain[mn]), lift<R>(s.mat_thickness[mn]), fr, rn_s, rn_h);
    const R kL = k * L;
    const R kappa = rho_p * rho_p / (sp * sp * sp) + rho * rho / (so * so * so);
    const R sg = w_sqrt(k * kappa * (float)(1.0 / kPi));
    const Cx<R> f1 = UtdTangent<R>::fresnel(sg * (g.length - tstar));
    const Cx<R> f0 = UtdTangent<R>::fresnel(sg * (-tstar));
    const Cx<R> dfr = f1 - f0;
    const Cx<R> dconj{dfr.re, -dfr.im};
    const Cx<R> trunc = cxc<R>(0.5f, 0.5f) * dconj;
    const R gamma = UtdTangent<R>::corner_blend(sg, tstar, g.length);
    const UtdBoundaries bd = utd_boundaries((double)val(g.n), (double)val(phi), (double)val(phi_p));
    const Cx<R> t1 = UtdTangent<R>::mend(UtdTangent<R>::term_n(g.n, phi - phi_p, +1, kL), !bd.incident_up, g.n, true,
                                         phi - phi_p, +1, 0, kL, trunc, gamma);
    const Cx<R> t2 = UtdTangent<R>::mend(UtdTangent<R>::term_n(g.n, phi - phi_p, -1, kL), bd.incident_up, g.n, true,
                                         phi - phi_p, -1, 1, kL, trunc, gamma);
    const Cx<R> t3 = UtdTangent<R>::mend(UtdTangent<R>::term_n(g.n, phi + phi_p, -1, kL), bd.reflect_0, g.n, true,
                                         phi + phi_p, -1, 2, kL, trunc, gamma);
    const Cx<R> t4 = UtdTangent<R>::mend(UtdTangent<R>::term_n(g.n, phi + phi_p, +1, kL), !bd.reflect_0, g.n, true,
                                         phi + phi_p, +1, 3, kL, trunc, gamma);
    const R pre_mag = lift<R>(1.0f) / (w_sqrt(k * (float)kTwoPi) * sin_b0 * (g.n * 2.0f));
    const Cx<R> pre = scale(Cx<R>{lift<R>(-(float)cos(-0.25 * kPi)), lift<R>(-(float)sin(-0.25 * kPi))}, pre_mag);
    const Cx<R> ds = pre * (t1 + t2 + r0_s * t3 + rn_s * t4);
    const Cx<R> dh = pre * (t1 + t2 + r0_h * t3 + rn_h * t4);
    const R amp = w_sqrt(sp / (so * (so + sp))) / (k * sp * 2.0f);
    const R ph = PhaseOf<R>::make(k, so + sp, k_d, length_d);
    const Cx<R> prop = scale(cexpj(ph), amp);
    arrival = sov;
    return c3_axis(b0_o, (ds * prop) * Cx<R>{-e_b, lift<R>(0.f)}) +
           c3_axis(phi_o, (dh * prop) * Cx<R>{-e_p, lift<R>(0.f)});
}

}  // namespace wc
