// Per-(transmitter, receiver) path discovery and field evaluation. One lane
// owns one cell; all lanes of a wavefront share the transmitter and walk the
// same candidate list (uniform, scalar-cache loads) while their receivers
// differ. Components: line of sight, image-method specular chains, first-order
// UTD edge diffraction and straight thin-sheet transmission. Every valid path
// is handed to a sink (accumulate / count / emit / tangent) in the cell's
// canonical order: depth, then component, then primitive sequence.
#pragma once

#include "wc_em.h"
#include "wc_scene.h"
#include "wc_trace.h"
#include "wc_utd.h"

namespace wc {

enum ComponentBits : int { kCompLos = 1, kCompRefl = 2, kCompDiff = 4, kCompTrans = 8 };
// wc_solve.components flag: a transmitter-receiver segment that crosses more
// walls than max_depth sets failure bit 4 instead of carrying no path
constexpr int kStrictPenetration = 16;
constexpr int kFailPenetration = 4;
enum ComponentId : int { kCidLos = 0, kCidRefl = 1, kCidDiff = 2, kCidTrans = 5 };

// accumulation slot of a component id: los 0, reflection 1, diffraction 2, transmission 3
__device__ __forceinline__ int slot_of_component(int component) {
    return component == kCidLos ? 0 : component == kCidRefl ? 1 : component == kCidDiff ? 2 : 3;
}

// minimum clearance of a visibility segment from its end points (m)
constexpr float kSegmentBias = 1.0e-4f;
// paths below this power are not emitted (diffraction existence floor)
constexpr float kGainFloor = 1.0e-30f;

struct TxSet {
    const float* pos;    // [T,3]
    const float* pol;    // [T,3]
    const float* power;  // [T]
    int count;
};

struct RxSet {
    const float* pos;         // receiver positions
    const float* pol;         // receiver polarizations, [1,3] (pol_stride 0) or per position
    int pol_stride;
    const int* cell_offset;   // [T+1]: cells of transmitter t are [cell_offset[t], cell_offset[t+1])
    int shared;               // 1: all transmitters share positions pos[cell - cell_offset[t]]
    int total_cells;
};

// Reflection candidates of every transmitter: chains of coplanar groups with
// their double-precision image points (image k = source mirrored by groups 0..k).
struct ChainSet {
    const int* offset;      // [T*kMaxDepth + 1]; chains of (t, depth d) at offset[t*kMaxDepth + d-1]..
    const int* groups;      // [K, kMaxDepth]
    const double* images;   // [K, kMaxDepth, 3]
};

struct PathRecord {
    int component;
    int depth;
    int candidate;                 // chain index, edge id, or -1
    int prim[kMaxDepth];           // face ids (reflection / transmission)
    int mat[kMaxDepth];
    float pos[kMaxDepth][3];
    float nrm[kMaxDepth][3];
    int edge;
    float length;                  // metres (float of the double length)
    double length_d;
    C3<float> field;               // unit-excitation complex 3-vector at the receiver
    Cx<float> coefficient;         // receiver projection of `field`
    Cx<float> path_field;          // coefficient * sqrt(tx power)
    float gain;                    // |path_field|^2
    float dep[3];                  // departure direction at the transmitter
    float arr[3];                  // arrival direction at the receiver (propagation direction)
};

__device__ __forceinline__ V3<float> ld3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ V3<double> to_d(const V3<float>& v) { return {(double)v.x, (double)v.y, (double)v.z}; }
__device__ __forceinline__ V3<float> to_f(const V3<double>& v) { return {(float)v.x, (float)v.y, (float)v.z}; }

__device__ __forceinline__ void finish_record(PathRecord& rec, const C3<float>& value, const V3<float>& arrival,
                                              const V3<float>& rx_pol, float amplitude_sqrt_power) {
    rec.field = value;
    rec.coefficient = project_receiver(value, arrival, rx_pol);
    rec.path_field = scale(rec.coefficient, amplitude_sqrt_power);
    rec.gain = abs2(rec.path_field);
    rec.arr[0] = arrival.x;
    rec.arr[1] = arrival.y;
    rec.arr[2] = arrival.z;
}

// ----------------------------------------------------------------- line of sight
// E = (p_tx - (p_tx.k)k) / (2 k d) e^{-jkd}; coefficient = p_rx,perp . E
// (reference native/channel/kernels/fields.cu:36-82).
__device__ inline C3<float> los_field(const V3<float>& dir, const V3<float>& tx_pol, float k, double k_d,
                                      double dist_d) {
    const float amp = 1.0f / (2.0f * fmaxf(k, kSmallEpsF) * fmaxf((float)dist_d, kEpsF));
    const float ph = neg_kd_phase(k_d, dist_d);
    const Cx<float> carrier = scale(cexpj(ph), amp);
    return c3_axis(transverse(dir, tx_pol), carrier);
}

// ----------------------------------------------------------------- reflection
__device__ __forceinline__ V3<double> group_normal(const SceneView& s, int g, double& offset) {
    const double* p = s.group_plane + 4 * g;
    offset = p[3];
    return {p[0], p[1], p[2]};
}

// First face of group g (ascending face id) whose triangle contains p.
__device__ inline int face_containing(const SceneView& s, int g, const V3<float>& p, const V3<float>& n) {
    const int begin = s.group_face_offset[g], end = s.group_face_offset[g + 1];
    for (int i = begin; i < end; ++i) {
        const int f = s.group_faces[i];
        const Tri t = s.faces[f];
        const V3<float> a{t.v0.x, t.v0.y, t.v0.z}, b{t.v1.x, t.v1.y, t.v1.z}, c{t.v2.x, t.v2.y, t.v2.z};
        const float e0 = dot(cross(b - a, p - a), n);
        const float e1 = dot(cross(c - b, p - b), n);
        const float e2 = dot(cross(a - c, p - c), n);
        if ((e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f) || (e0 <= 0.0f && e1 <= 0.0f && e2 <= 0.0f)) return f;
    }
    return -1;
}

// Visibility of the D+1 legs of a reflection chain through its hit points
// (each leg ignores the groups it starts and ends on). The legs are tested
// from the receiver back to the transmitter: the verdict is the AND of the
// legs, and a receiver near the ground is the end most often blocked, so the
// first walk most often ends the test (scaled city, chain validity pass per
// batch: 50.0 ms transmitter leg first, 33.3 ms receiver leg first, 33.9 ms
// receiver leg, then transmitter leg, then the middle legs).
template <int D, bool kTop = false>
__device__ inline bool chain_legs_clear(const SceneView& s, const int* groups, const V3<double>* hits,
                                        const V3<float>& tx, const V3<float>& rx) {
#pragma unroll
    for (int j = 0; j <= D; ++j) {
        const int k = D - j;  // leg k runs from point k to point k + 1
        const V3<float> a = k == 0 ? tx : to_f(hits[k - 1]);
        const V3<float> b = k < D ? to_f(hits[k]) : rx;
        const int ga = k == 0 ? -1 : groups[k - 1];
        const int gb = k < D ? groups[k] : -1;
        const V3<float> seg = b - a;
        const float len = norm(seg);
        if (len > 2.0f * kSegmentBias) {
            const float eps = kSegmentBias / len;
            const RayPre r = ray_setup(a.x, a.y, a.z, seg.x, seg.y, seg.z);
            if (occluded<kTop>(s, r, eps, 1.0f - eps, ga, gb)) return false;
        }
    }
    return true;
}

// Validates one image-method chain for receiver rx: back-projects from the
// receiver through the images, requires each hit strictly between its image
// and the next point and inside a face of its group, then (kLegs) tests every
// leg for occlusion (legs skip the groups they start and end on). On success
// fills hits/faces and returns true. The depth D is a compile-time constant
// so the per-bounce arrays stay in registers.
template <int D, bool kLegs = true, bool kTop = false>
__device__ inline bool solve_chain(const SceneView& s, const int* groups, const double* images, const V3<float>& tx,
                                   const V3<float>& rx, V3<double>* hits, int* faces) {
    V3<double> next = to_d(rx);
#pragma unroll
    for (int k = D - 1; k >= 0; --k) {
        const int g = groups[k];
        double off;
        const V3<double> n = group_normal(s, g, off);
        const V3<double> img{images[3 * k], images[3 * k + 1], images[3 * k + 2]};
        const double si = dot(n, img) - off;
        const double sn = dot(n, next) - off;
        if (!(si * sn < 0.0)) return false;
        const double t = si / (si - sn);
        const V3<double> h = img + (next - img) * t;
        const int f = face_containing(s, g, to_f(h), to_f(n));
        if (f < 0) return false;
        hits[k] = h;
        faces[k] = f;
        next = h;
    }
    if (!kLegs) return true;  // leg visibility already established (chain masks)
    return chain_legs_clear<D, kTop>(s, groups, hits, tx, rx);
}

// Field of a reflection chain with the interaction points fixed: the Jones
// transport of fields.cu:113-204 with amplitude 1/(2kL) and phase -kL, L the
// double-precision unfolded length.
template <int D>
__device__ inline C3<float> reflection_field(const SceneView& s, const V3<float>& tx, const V3<float>& rx,
                                             const V3<float>* hits, const V3<float>* normals, const int* mats,
                                             const V3<float>& tx_pol, float freq, double k_d, double length_d,
                                             V3<float>& arrival, V3<float>& departure) {
    const V3<float> ez = v3c<float>(0.f, 0.f, 1.f);
    V3<float> prev = tx;
    V3<float> incident = safe_normalize(hits[0] - prev, ez);
    departure = incident;
    C3<float> value = c3_from_real(transverse(incident, tx_pol));
    V3<float> outgoing = incident;
#pragma unroll
    for (int b = 0; b < D; ++b) {
        incident = safe_normalize(hits[b] - prev, outgoing);
        const int m = mats[b];
        value = reflect_c3(value, incident, normals[b], s.mat_eps[m], s.mat_sigma[m], s.mat_mu[m], s.mat_gain[m],
                           s.mat_thickness[m], freq, outgoing);
        prev = hits[b];
    }
    arrival = safe_normalize(rx - prev, outgoing);
    const float k = (float)k_d;
    const float amp = 1.0f / (2.0f * k * fmaxf((float)length_d, kEpsF));
    const Cx<float> prop = scale(cexpj(neg_kd_phase(k_d, length_d)), amp);
    return c3_scale(value, prop);
}

// ----------------------------------------------------------------- transmission
// Straight thin-sheet chain: per wall the Rouard stack t_TE / t_TM on the local
// s/p basis; carrier length L - sum(thickness cos theta); amplitude 1/(2kL)
// (reference tests/ad/_reference_fields.py:269-336).
__device__ inline C3<float> transmission_field(const SceneView& s, int count, const Hit* hits, const V3<float>& dir,
                                               const V3<float>& tx_pol, float freq, double k_d, double length_d,
                                               int* mats_out, V3<float>* normals_out) {
    const V3<float> ez = v3c<float>(0.f, 0.f, 1.f);
    C3<float> value = c3_from_real(transverse(dir, tx_pol));
    double carrier = length_d;
    for (int w = 0; w < count; ++w) {
        const int g = hits[w].group;
        double off;
        const V3<double> nd = group_normal(s, g, off);
        V3<float> n = safe_normalize(to_f(nd), ez);
        if (dot(dir, n) > 0.0f) n = -n;
        const float ct = fminf(fmaxf(fabsf(dot(dir, n)), kSmallEpsF), 1.0f);
        const int m = s.face_material[hits[w].face];
        const int lo = s.mat_layer_offset[m], lc = s.mat_layer_count[m];
        const Cx<float> t_te = stack_t<float>(ct, s.layer_thickness + lo, s.layer_eps + lo, s.layer_sigma + lo,
                                              s.layer_mu + lo, lc, freq, 0, nullptr, nullptr);
        const Cx<float> t_tm = stack_t<float>(ct, s.layer_thickness + lo, s.layer_eps + lo, s.layer_sigma + lo,
                                              s.layer_mu + lo, lc, freq, 1, nullptr, nullptr);
        const V3<float> sax = safe_normalize(cross(n, dir), stable_perp_basis(dir, n));
        const V3<float> pax = safe_normalize(cross(sax, dir), stable_perp_basis(dir, sax));
        const Cx<float> e_s = c3_dot_real(value, sax);
        const Cx<float> e_p = c3_dot_real(value, pax);
        value = c3_axis(sax, t_te * e_s) + c3_axis(pax, t_tm * e_p);
        double wall = 0.0;
        for (int l = 0; l < lc; ++l) wall += (double)fmaxf(s.layer_thickness[lo + l], 0.0f);
        carrier -= wall * (double)ct;
        mats_out[w] = m;
        normals_out[w] = n;
    }
    const float k = (float)k_d;
    const float amp = 1.0f / (2.0f * k * fmaxf((float)length_d, kEpsF));
    const Cx<float> prop = scale(cexpj(neg_kd_phase(k_d, carrier)), amp);
    return c3_scale(value, prop);
}

// ----------------------------------------------------------------- diffraction
struct DiffGeometry {
    V3<double> q;       // stationary (Keller) point on the infinite edge line
    double tstar;       // its parameter along the edge
    double sp, s;       // |T - Q|, |R - Q|
    double phi_p, phi;  // incidence / observation angles in [0, n pi]
    double rho_p, rho;  // distances of T and R from the edge line
};

__device__ __forceinline__ double wedge_angle(const V3<double>& u, const V3<double>& t0, const V3<double>& n0) {
    double a = atan2(dot(u, n0), dot(u, t0));
    if (a < 0.0) a += kTwoPi;
    return a;
}

// Keller point by unfolding: t* = a + (b - a) rho_T / (rho_T + rho_R).
// With observation = false the receiver's wedge angle phi (used only by the
// field) is left unset; the existence test and every other member are the
// same.
// (diffraction_geometry_d takes double end points: image sources and
// receivers of the coupled paths keep their full precision.)
__device__ inline bool diffraction_geometry_d(const Edge& e, const V3<double>& T, const V3<double>& Rr,
                                              DiffGeometry& g, bool observation = true) {
    const V3<double> A{e.ax, e.ay, e.az}, d{e.dx, e.dy, e.dz};
    const double a = dot(T - A, d), b = dot(Rr - A, d);
    const V3<double> pt = T - A - d * a, pr = Rr - A - d * b;
    g.rho_p = norm(pt);
    g.rho = norm(pr);
    if (g.rho_p < 1.0e-6 || g.rho < 1.0e-6) return false;
    g.tstar = a + (b - a) * g.rho_p / (g.rho_p + g.rho);
    g.q = A + d * g.tstar;
    g.sp = norm(T - g.q);
    g.s = norm(Rr - g.q);
    const V3<double> t0{e.t0x, e.t0y, e.t0z}, n0{e.n0x, e.n0y, e.n0z};
    const double npi = (double)e.n * kPi;
    g.phi_p = wedge_angle(pt, t0, n0);
    if (g.phi_p >= npi) return false;  // source inside the wedge
    if (!observation) return true;
    double phi = wedge_angle(pr, t0, n0);
    if (phi > npi) phi = (phi - npi < kTwoPi - phi) ? npi : 0.0;  // nearest face of the exterior domain
    g.phi = phi;
    return true;
}

__device__ inline bool diffraction_geometry(const Edge& e, const V3<float>& tx, const V3<float>& rx,
                                            DiffGeometry& g, bool observation = true) {
    return diffraction_geometry_d(e, to_d(tx), to_d(rx), g, observation);
}

// Diffracted complex 3-vector at the receiver (unit excitation), K-P dyadic
// E^d = [-D_s b0 b0' - D_h phi phi'] . E^i(Q) sqrt(s'/(s (s+s'))) e^{-jks},
// E^i(Q) = (p_tx - (p_tx.s')s') e^{-jks'} / (2ks'), with the finite-edge
// truncation inside the coefficients (utd_wedge). Face reflection coefficients (Luebbers) are the slab
// coefficients of each face at its grazing angle: soft <- TE, hard <- TM.
__device__ inline C3<float> diffraction_field(const SceneView& s, const Edge& e, const DiffGeometry& g,
                                              const V3<float>& tx, const V3<float>& rx, const V3<float>& tx_pol,
                                              float freq, double k_d, V3<float>& arrival, V3<float>& departure) {
    const V3<double> ed{e.dx, e.dy, e.dz};
    const V3<double> T = to_d(tx), Rr = to_d(rx);
    const V3<double> si = (g.q - T) * (1.0 / g.sp);   // incident direction s'
    const V3<double> so = (Rr - g.q) * (1.0 / g.s);   // diffracted direction s
    const double cb = dot(si, ed);
    const double sin_b0 = sqrt(fmax(1.0 - cb * cb, 0.0));
    V3<double> phi_i = -cross(ed, si);
    phi_i = phi_i * (1.0 / fmax(norm(phi_i), 1.0e-30));
    const V3<double> b0_i = cross(phi_i, si);
    V3<double> phi_o = cross(ed, so);
    phi_o = phi_o * (1.0 / fmax(norm(phi_o), 1.0e-30));
    const V3<double> b0_o = cross(phi_o, so);
    const V3<float> sif = to_f(si), sof = to_f(so);
    const V3<float> tpol = transverse(sif, tx_pol);
    const float e_b = dot(tpol, to_f(b0_i));
    const float e_p = dot(tpol, to_f(phi_i));
    const double L = g.s * g.sp * sin_b0 * sin_b0 / (g.s + g.sp);
    const double npi = (double)e.n * kPi;
    const int m0 = e.material0;
    const int mn = e.material1 >= 0 ? e.material1 : e.material0;
    Cx<float> r0_s, r0_h, rn_s, rn_h;
    slab_fresnel<float>((float)fabs(sin(g.phi_p)), s.mat_eps[m0], s.mat_sigma[m0], s.mat_mu[m0], s.mat_gain[m0],
                        s.mat_thickness[m0], freq, r0_s, r0_h);
    slab_fresnel<float>((float)fabs(sin(npi - g.phi)), s.mat_eps[mn], s.mat_sigma[mn], s.mat_mu[mn],
                        s.mat_gain[mn], s.mat_thickness[mn], freq, rn_s, rn_h);
    // finite-edge truncation factor and corner blend, folded into the terms
    const double kappa = g.rho_p * g.rho_p / (g.sp * g.sp * g.sp) + g.rho * g.rho / (g.s * g.s * g.s);
    const double sigma = sqrt(k_d * kappa / kPi);
    double f1r, f1i, f0r, f0i;
    fresnel_complex(sigma * ((double)e.length - g.tstar), f1r, f1i);
    fresnel_complex(sigma * (0.0 - g.tstar), f0r, f0i);
    const double dr = f1r - f0r, di = -(f1i - f0i);  // conj(difference)
    const double u = sigma * fmax(-g.tstar, g.tstar - (double)e.length);
    const FiniteEdge fe{0.5 * dr - 0.5 * di, 0.5 * di + 0.5 * dr, 1.0 / (1.0 + exp(u * 10.0))};
    const UtdCoefficients D = utd_wedge((double)e.n, g.phi, g.phi_p, sin_b0, k_d, L, r0_s, rn_s, r0_h, rn_h, fe);
    // spreading and incident amplitude: sqrt(s'/(s(s+s'))) / (2 k s')
    const double amp = sqrt(g.sp / (g.s * (g.s + g.sp))) / (2.0 * k_d * g.sp);
    // total propagation scalar: amp * e^{-jk(s+s')}
    const float ph = neg_kd_phase(k_d, g.s + g.sp);
    const Cx<float> prop = Cx<float>{(float)amp, 0.0f} * cexpj(ph);
    const Cx<float> ds{D.ds_re, D.ds_im}, dh{D.dh_re, D.dh_im};
    const Cx<float> cb_o = scale(ds, -e_b) * prop;
    const Cx<float> cp_o = scale(dh, -e_p) * prop;
    arrival = sof;
    departure = sif;
    return c3_axis(to_f(b0_o), cb_o) + c3_axis(to_f(phi_o), cp_o);
}

}  // namespace wc
