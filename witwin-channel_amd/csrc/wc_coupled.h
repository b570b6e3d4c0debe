// Coupled reflection-diffraction paths, one candidate at a time (reference
// witwin/channel/interactions/coupled.py:93-230 candidate order,
// native/channel/kernels/coupled.cu:2279-2710 geometry): R->D through the
// image of the transmitter, D->R through the image of the receiver, D->D by a
// bisection of the path length's derivative along edge 2 (edge-1 Keller
// point in closed form); every leg tested through the 4-wide tree; fields
// composed from the thin-slab Jones reflection and the first-order UTD wedge
// coefficient (K-P, Luebbers face coefficients, E1d finite-edge assembly).
// Device code; also built as host C++ for the CPU checks.
#pragma once

#include "wc_paths.h"

namespace wc {

constexpr double kCoupledEps = 1.0e-6;   // interior margin on edges and legs (reference coupled.cu:2269)
constexpr int kDdBisection = 48;         // double-precision bisection steps of the D->D stationary point

struct CoupledRecord {
    float pf_re, pf_im, gain, kind;  // kind: 0 none, 3 R->D, 4 D->R, 7 D->D (the component id)
};

// Any triangle on the segment a-b except those of the up-to-four groups the
// segment starts or ends on.
__device__ inline bool leg_blocked(const SceneView& s, const V3<double>& a_d, const V3<double>& b_d, int g0, int g1,
                                   int g2, int g3) {
    const V3<float> a = to_f(a_d), b = to_f(b_d);
    const V3<float> seg = b - a;
    const float len = norm(seg);
    if (len <= 2.0f * kSegmentBias || s.tri_count == 0) return false;
    const float eps = kSegmentBias / len;
    const RayPre r = ray_setup(a.x, a.y, a.z, seg.x, seg.y, seg.z);
    bool blocked = false;
    walk_wide(s, r, eps, 1.0f - eps, [&](const Tri& tri) {
        const int g = bits_i(tri.v0.w);
        blocked = g != g0 && g != g1 && g != g2 && g != g3 && tri_hit(r, tri, eps, 1.0f - eps) >= 0.0f;
        return blocked;
    });
    return blocked;
}

__device__ __forceinline__ Cx<float> cmul(const Cx<float>& a, const Cx<float>& b) { return a * b; }

// UTD diffraction of the incident field einc (complex vector at the Keller
// point, transverse to the incidence direction, spreading from its source
// already applied) arriving from src at distance g.sp and observed at obs at
// distance g.s: E^d = [-D_s b0 b0' - D_h phi phi'] . einc sqrt(s'/(s(s+s')))
// e^{-jks}, the coefficients of diffraction_field (K-P with Luebbers face
// coefficients and the E1d finite-edge assembly).
__device__ inline C3<float> wedge_diffract(const SceneView& s, const Edge& e, const DiffGeometry& g,
                                           const V3<double>& src, const V3<double>& obs, const C3<float>& einc,
                                           float freq, double k_d, V3<double>& so) {
    const V3<double> ed{e.dx, e.dy, e.dz};
    const V3<double> si = (g.q - src) * (1.0 / g.sp);
    so = (obs - g.q) * (1.0 / g.s);
    const double cb = dot(si, ed);
    const double sin_b0 = sqrt(fmax(1.0 - cb * cb, 0.0));
    V3<double> phi_i = -cross(ed, si);
    phi_i = phi_i * (1.0 / fmax(norm(phi_i), 1.0e-30));
    const V3<double> b0_i = cross(phi_i, si);
    V3<double> phi_o = cross(ed, so);
    phi_o = phi_o * (1.0 / fmax(norm(phi_o), 1.0e-30));
    const V3<double> b0_o = cross(phi_o, so);
    const Cx<float> e_b = c3_dot_real(einc, to_f(b0_i));
    const Cx<float> e_p = c3_dot_real(einc, to_f(phi_i));
    const double L = g.s * g.sp * sin_b0 * sin_b0 / (g.s + g.sp);
    const double npi = (double)e.n * kPi;
    const int m0 = e.material0;
    const int mn = e.material1 >= 0 ? e.material1 : e.material0;
    Cx<float> r0_s, r0_h, rn_s, rn_h;
    slab_fresnel<float>((float)fabs(sin(g.phi_p)), s.mat_eps[m0], s.mat_sigma[m0], s.mat_mu[m0], s.mat_gain[m0],
                        s.mat_thickness[m0], freq, r0_s, r0_h);
    slab_fresnel<float>((float)fabs(sin(npi - g.phi)), s.mat_eps[mn], s.mat_sigma[mn], s.mat_mu[mn],
                        s.mat_gain[mn], s.mat_thickness[mn], freq, rn_s, rn_h);
    const double kappa = g.rho_p * g.rho_p / (g.sp * g.sp * g.sp) + g.rho * g.rho / (g.s * g.s * g.s);
    const double sigma = sqrt(k_d * kappa / kPi);
    double f1r, f1i, f0r, f0i;
    fresnel_complex(sigma * ((double)e.length - g.tstar), f1r, f1i);
    fresnel_complex(sigma * (0.0 - g.tstar), f0r, f0i);
    const double dr = f1r - f0r, di = -(f1i - f0i);
    const double u = sigma * fmax(-g.tstar, g.tstar - (double)e.length);
    const FiniteEdge fe{0.5 * dr - 0.5 * di, 0.5 * di + 0.5 * dr, 1.0 / (1.0 + exp(u * 10.0))};
    const UtdCoefficients D = utd_wedge((double)e.n, g.phi, g.phi_p, sin_b0, k_d, L, r0_s, rn_s, r0_h, rn_h, fe);
    const double amp = sqrt(g.sp / (g.s * (g.s + g.sp)));
    const Cx<float> prop = scale(cexpj(neg_kd_phase(k_d, g.s)), (float)amp);
    const Cx<float> ds{D.ds_re, D.ds_im}, dh{D.dh_re, D.dh_im};
    const Cx<float> cb_o = cmul(cmul(ds, scale(e_b, -1.0f)), prop);
    const Cx<float> cp_o = cmul(cmul(dh, scale(e_p, -1.0f)), prop);
    return c3_axis(to_f(b0_o), cb_o) + c3_axis(to_f(phi_o), cp_o);
}

// spherical incident field p_perp e^{-jks'} / (2ks') of the transmitter at the
// Keller point g.q (incidence along src -> q)
__device__ inline C3<float> spherical_incident(const V3<double>& src, const DiffGeometry& g, const V3<float>& pol,
                                               double k_d) {
    const V3<float> si = to_f((g.q - src) * (1.0 / g.sp));
    const Cx<float> c = scale(cexpj(neg_kd_phase(k_d, g.sp)), (float)(1.0 / (2.0 * k_d * g.sp)));
    return c3_axis(transverse(si, pol), c);
}

__device__ __forceinline__ bool on_edge(const Edge& e, double t) {
    return t > kCoupledEps && t < (double)e.length - kCoupledEps;
}

__device__ __forceinline__ V3<double> mirror(const V3<double>& p, const V3<double>& n, double off) {
    return p - n * (2.0 * (dot(n, p) - off));
}

// reflection at R of group grp: resolves the face, returns its material (or -1)
__device__ __forceinline__ int reflection_face(const SceneView& s, int grp, const V3<double>& R, const V3<double>& n) {
    const int f = face_containing(s, grp, to_f(R), to_f(n));
    return f < 0 ? -1 : s.face_material[f];
}

// TX -> reflect(group) -> diffract(edge) -> RX (component 3)
__device__ inline bool coupled_rd(const SceneView& s, int grp, const Edge& e, const V3<double>& T, const V3<double>& X,
                           const V3<float>& tx_pol, float freq, double k_d, C3<float>& field, V3<double>& arrival) {
    double off;
    const V3<double> n = group_normal(s, grp, off);
    const V3<double> I = mirror(T, n, off);
    DiffGeometry g;
    if (!diffraction_geometry_d(e, I, X, g) || !on_edge(e, g.tstar)) return false;
    const double den = dot(n, g.q - I);
    if (fabs(den) < 1.0e-12) return false;
    const double t = (off - dot(n, I)) / den;
    if (!(t > kCoupledEps && t < 1.0 - kCoupledEps)) return false;
    const V3<double> R = I + (g.q - I) * t;
    const int m = reflection_face(s, grp, R, n);
    if (m < 0) return false;
    if (leg_blocked(s, T, R, grp, -1, -1, -1) || leg_blocked(s, R, g.q, grp, e.group0, e.group1, -1) ||
        leg_blocked(s, g.q, X, e.group0, e.group1, -1, -1))
        return false;
    const V3<float> ez = v3c<float>(0.f, 0.f, 1.f);
    const V3<float> d1 = safe_normalize(to_f(R - T), ez);
    V3<float> outgoing;
    C3<float> value = c3_from_real(transverse(d1, tx_pol));
    value = reflect_c3(value, d1, to_f(n), s.mat_eps[m], s.mat_sigma[m], s.mat_mu[m], s.mat_gain[m],
                       s.mat_thickness[m], freq, outgoing);
    const Cx<float> c = scale(cexpj(neg_kd_phase(k_d, g.sp)), (float)(1.0 / (2.0 * k_d * g.sp)));
    field = wedge_diffract(s, e, g, I, X, c3_scale(value, c), freq, k_d, arrival);
    return true;
}

// TX -> diffract(edge) -> reflect(group) -> RX (component 4)
__device__ inline bool coupled_dr(const SceneView& s, int grp, const Edge& e, const V3<double>& T, const V3<double>& X,
                           const V3<float>& tx_pol, float freq, double k_d, C3<float>& field, V3<double>& arrival) {
    double off;
    const V3<double> n = group_normal(s, grp, off);
    const V3<double> Xi = mirror(X, n, off);
    DiffGeometry g;
    if (!diffraction_geometry_d(e, T, Xi, g) || !on_edge(e, g.tstar)) return false;
    const double den = dot(n, Xi - g.q);
    if (fabs(den) < 1.0e-12) return false;
    const double t = (off - dot(n, g.q)) / den;
    if (!(t > kCoupledEps && t < 1.0 - kCoupledEps)) return false;
    const V3<double> R = g.q + (Xi - g.q) * t;
    const int m = reflection_face(s, grp, R, n);
    if (m < 0) return false;
    if (leg_blocked(s, T, g.q, e.group0, e.group1, -1, -1) || leg_blocked(s, g.q, R, e.group0, e.group1, grp, -1) ||
        leg_blocked(s, R, X, grp, -1, -1, -1))
        return false;
    V3<double> so;
    const C3<float> at_r = wedge_diffract(s, e, g, T, Xi, spherical_incident(T, g, tx_pol, k_d), freq, k_d, so);
    const V3<float> ez = v3c<float>(0.f, 0.f, 1.f);
    const V3<float> inc = safe_normalize(to_f(R - g.q), ez);
    V3<float> outgoing;
    field = reflect_c3(at_r, inc, to_f(n), s.mat_eps[m], s.mat_sigma[m], s.mat_mu[m], s.mat_gain[m],
                       s.mat_thickness[m], freq, outgoing);
    arrival = safe_normalize(X - R, to_d(outgoing));
    return true;
}

// Keller point on the line of edge e for the end points a and b (closed form
// of the unfolding; reference coupled.cu:2424-2470 for its use in D->D)
__device__ __forceinline__ double keller_t(const Edge& e, const V3<double>& a, const V3<double>& b) {
    const V3<double> A{e.ax, e.ay, e.az}, d{e.dx, e.dy, e.dz};
    const double ta = dot(a - A, d), tb = dot(b - A, d);
    const double ra = norm(a - A - d * ta), rb = norm(b - A - d * tb);
    const double w = ra + rb;
    return w > 0.0 ? ta + (tb - ta) * ra / w : ta;
}

// TX -> diffract(e1) -> diffract(e2) -> RX (component 7): the path length
// over Q2 = A2 + d2 t2 with Q1 the Keller point of (TX, Q2) on edge 1 is
// convex in t2; its derivative d2 . (unit(Q2 - Q1) + unit(Q2 - RX)) is
// bracketed on the edge interior and bisected.
__device__ inline bool coupled_dd(const SceneView& s, const Edge& e1, const Edge& e2, const V3<double>& T,
                           const V3<double>& X, const V3<float>& tx_pol, float freq, double k_d, C3<float>& field,
                           V3<double>& arrival) {
    const V3<double> A1{e1.ax, e1.ay, e1.az}, d1{e1.dx, e1.dy, e1.dz};
    const V3<double> A2{e2.ax, e2.ay, e2.az}, d2{e2.dx, e2.dy, e2.dz};
    // collinear edges (one physical line) have no stationary middle leg, and
    // two edges of one face group would join by a leg grazing that face
    // (incidence angle 0 on the second wedge, decided by the sign of a
    // rounding error): both are excluded
    const V3<double> c = cross(d1, d2);
    if (norm(c) < 1.0e-9 && norm(cross(A2 - A1, d1)) < 1.0e-6) return false;
    const bool shared = (e1.group0 >= 0 && (e1.group0 == e2.group0 || e1.group0 == e2.group1)) ||
                        (e1.group1 >= 0 && (e1.group1 == e2.group0 || e1.group1 == e2.group1));
    if (shared) return false;
    auto slope = [&](double t2, double& t1) -> double {
        const V3<double> q2 = A2 + d2 * t2;
        t1 = keller_t(e1, T, q2);
        const V3<double> q1 = A1 + d1 * t1;
        const V3<double> mid = q2 - q1, out = q2 - X;
        const double lm = norm(mid), lo = norm(out);
        if (!(lm > kCoupledEps) || !(lo > kCoupledEps)) return __builtin_nan("");
        return dot(d2, mid * (1.0 / lm) + out * (1.0 / lo));
    };
    double lo = kCoupledEps, hi = (double)e2.length - kCoupledEps, t1;
    if (!(hi > lo)) return false;
    const double s_lo = slope(lo, t1), s_hi = slope(hi, t1);
    if (!(s_lo < 0.0 && s_hi > 0.0)) return false;
    for (int it = 0; it < kDdBisection; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double sm = slope(mid, t1);
        if (!(sm == sm)) return false;
        if (sm > 0.0) hi = mid; else lo = mid;
    }
    const double t2 = 0.5 * (lo + hi);
    slope(t2, t1);
    if (!on_edge(e1, t1)) return false;
    const V3<double> Q1 = A1 + d1 * t1, Q2 = A2 + d2 * t2;
    DiffGeometry g1, g2;
    if (!diffraction_geometry_d(e1, T, Q2, g1) || !diffraction_geometry_d(e2, Q1, X, g2)) return false;
    g2.q = Q2;
    g2.tstar = t2;
    g2.sp = norm(Q2 - Q1);
    g2.s = norm(X - Q2);
    if (leg_blocked(s, T, Q1, e1.group0, e1.group1, -1, -1) ||
        leg_blocked(s, Q1, Q2, e1.group0, e1.group1, e2.group0, e2.group1) ||
        leg_blocked(s, Q2, X, e2.group0, e2.group1, -1, -1))
        return false;
    V3<double> so;
    const C3<float> at_q2 = wedge_diffract(s, e1, g1, T, Q2, spherical_incident(T, g1, tx_pol, k_d), freq, k_d, so);
    field = wedge_diffract(s, e2, g2, Q1, X, at_q2, freq, k_d, arrival);
    return true;
}


// Candidate k of a (transmitter, receiver) pair: k < G*E R->D over (k / E,
// k % E); then G*E D->R; then the ordered edge pairs (e1, e2 != e1) of D->D.
// kinds: 1 R->D, 2 D->R, 4 D->D. Returns the path field sqrt(P) x
// receiver-projected coefficient, its gain and the component id (0: none).
__device__ inline CoupledRecord coupled_candidate(const SceneView& s, int k, int kinds, const V3<double>& T,
                                                  const V3<double>& X, const V3<float>& tx_pol,
                                                  const V3<float>& rx_pol, float sqrt_power, float freq,
                                                  double k_d) {
    const int G = s.group_count, E = s.edge_count;
    const int ge = G * E;
    C3<float> field;
    V3<double> arrival;
    bool ok = false;
    float kind = 0.0f;
    if (k < ge) {
        kind = 3.0f;
        if (kinds & 1) ok = coupled_rd(s, k / E, s.edges[k % E], T, X, tx_pol, freq, k_d, field, arrival);
    } else if (k < 2 * ge) {
        kind = 4.0f;
        const int j = k - ge;
        if (kinds & 2) ok = coupled_dr(s, j / E, s.edges[j % E], T, X, tx_pol, freq, k_d, field, arrival);
    } else if (kinds & 4) {
        kind = 7.0f;
        const int j = k - 2 * ge;
        const int e1 = j / (E - 1), r = j % (E - 1);
        const int e2 = r < e1 ? r : r + 1;
        ok = coupled_dd(s, s.edges[e1], s.edges[e2], T, X, tx_pol, freq, k_d, field, arrival);
    }
    CoupledRecord out{0.0f, 0.0f, 0.0f, 0.0f};
    if (ok) {
        const Cx<float> coef = project_receiver(field, to_f(arrival), rx_pol);
        const Cx<float> pf = scale(coef, sqrt_power);
        const float gain = pf.re * pf.re + pf.im * pf.im;
        if (gain >= kGainFloor && gain == gain) out = {pf.re, pf.im, gain, kind};
    }
    return out;
}

}  // namespace wc
