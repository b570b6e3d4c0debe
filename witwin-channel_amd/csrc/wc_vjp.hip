// Fixed-topology adjoint of the fused radiomap.
//
// For every cell the forward pass stored its path descriptors (component,
// candidate). The adjoint recomputes each stored path from the transmitter
// position, the frequency and its per-interaction material parameters with
// dual numbers (no visibility traversal: the topology is frozen; only a
// transmission path re-gathers its walls), contracts the path tangents with
// the cell adjoint of the coherent sum
//   dL/dF_p = gP * 2 S + gS + gP_slot * 2 C_slot + gC_slot   (coherent)
// and reduces the per-lane gradients over each 64-cell tile with a fixed
// butterfly, writing one partial vector per tile. Summing the partials over
// tiles (in tile order) gives dL/d(eps_r, sigma) per material, dL/df and
// dL/d(tx position) per transmitter.
#include "wc_api_internal.h"
#include "wc_tangent.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

// tangent directions carried per dual-number pass (a path with more local
// parameters takes several passes)
#ifndef WC_VJP_TANGENTS
#define WC_VJP_TANGENTS 8
#endif
constexpr int kTangents = WC_VJP_TANGENTS;
constexpr int kMaxMaterials = 16;
using D8 = Dual<kTangents>;

__device__ __forceinline__ Cx<float> ld_c(const float* p, size_t i) { return {p[2 * i], p[2 * i + 1]}; }

struct CellAdjoint {
    float a_re[4], a_im[4];   // coherent per-slot field adjoint
    float g_inc[4];           // incoherent: dL/dgain of each slot (gP + gP_slot)
    float g_f_inc;            // incoherent: gF_re / sqrt(P)
    Cx<float> b[3];           // adjoint of the diffraction vector sum (per axis)
    bool coherent;
};

__device__ CellAdjoint cell_adjoint(const wc_adjoint& g, const wc_solve& q, int cell) {
    const int cells = q.total_cells;
    CellAdjoint a;
    a.coherent = q.coherent != 0;
    const float gp = g.grad_power_total ? g.grad_power_total[cell] : 0.0f;
    const Cx<float> gf = g.grad_field_total ? ld_c(g.grad_field_total, cell) : Cx<float>{0.f, 0.f};
    const Cx<float> S = ld_c(g.field_total, cell);
    for (int s = 0; s < 4; ++s) {
        const float gps = g.grad_component_power ? g.grad_component_power[(size_t)s * cells + cell] : 0.0f;
        const Cx<float> gcf =
            g.grad_component_field ? ld_c(g.grad_component_field, (size_t)s * cells + cell) : Cx<float>{0.f, 0.f};
        const Cx<float> C = ld_c(g.component_field, (size_t)s * cells + cell);
        a.a_re[s] = gp * 2.0f * S.re + gf.re + gps * 2.0f * C.re + gcf.re;
        a.a_im[s] = gp * 2.0f * S.im + gf.im + gps * 2.0f * C.im + gcf.im;
        a.g_inc[s] = gp + gps;
    }
    const float p = g.power_total ? g.power_total[cell] : 0.0f;
    a.g_f_inc = p > 0.0f ? gf.re / sqrtf(p) : 0.0f;
    const float gd = g.grad_diffraction_power ? g.grad_diffraction_power[cell] : 0.0f;
    for (int j = 0; j < 3; ++j) {
        const Cx<float> v = g.diffraction_vector ? ld_c(g.diffraction_vector, (size_t)3 * cell + j)
                                                 : Cx<float>{0.f, 0.f};
        a.b[j] = {gd * 2.0f * v.re, gd * 2.0f * v.im};
    }
    return a;
}

struct LaneGrad {
    float eps[kMaxMaterials];
    float sig[kMaxMaterials];
    float freq;
    float tx[3];
};

__device__ __forceinline__ void add_material(LaneGrad& g, int m, float de, float ds) {
#pragma unroll
    for (int i = 0; i < kMaxMaterials; ++i) {
        if (i == m) {
            g.eps[i] += de;
            g.sig[i] += ds;
        }
    }
}

// Local parameter slots of one path: 0..2 tx xyz, 3 frequency, 4+2k eps of
// interaction k, 5+2k sigma of interaction k.
struct PathParams {
    int count;               // 4 + 2 * interactions
    int material[kMaxDepth];
    float eps[kMaxDepth];
    float sig[kMaxDepth];
};

__device__ __forceinline__ D8 seeded(float v, int slot, int pass) {
    const int local = slot - pass * kTangents;
    return seed<kTangents>(v, (local >= 0 && local < kTangents) ? local : -1);
}

// d(contribution)/d(slot) for the seeded slots of one pass
template <int N>
__device__ __forceinline__ void contract(const C3<Dual<N>>& field, const V3<Dual<N>>& arrival,
                                         const V3<float>& rx_pol, float sqrt_power, int slot_id,
                                         const CellAdjoint& adj, bool is_diff, float out[N]) {
    using DN = Dual<N>;
    const Cx<DN> coef = project_receiver(field, arrival, lift3<DN>(rx_pol));
    float ar, ai;
    if (adj.coherent) {
        ar = adj.a_re[slot_id];
        ai = adj.a_im[slot_id];
    } else {
        // gain = |F|^2 with F = coef * sqrt_power
        const float fr = coef.re.v * sqrt_power, fi = coef.im.v * sqrt_power;
        ar = adj.g_inc[slot_id] * 2.0f * fr + adj.g_f_inc * fr;
        ai = adj.g_inc[slot_id] * 2.0f * fi + adj.g_f_inc * fi;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        float v = (ar * coef.re.d[i] + ai * coef.im.d[i]) * sqrt_power;
        if (is_diff) {
            const Cx<DN> comps[3] = {field.x, field.y, field.z};
            for (int j = 0; j < 3; ++j)
                v += (adj.b[j].re * comps[j].re.d[i] + adj.b[j].im * comps[j].im.d[i]) * sqrt_power;
        }
        out[i] = v;
    }
}

// dL/d(coefficient) of one path as (re, im) weights
__device__ __forceinline__ void coef_weights(const Cx<float>& coef, float sqrt_power, int slot_id,
                                             const CellAdjoint& adj, float& ar, float& ai) {
    if (adj.coherent) {
        ar = adj.a_re[slot_id];
        ai = adj.a_im[slot_id];
    } else {
        const float fr = coef.re * sqrt_power, fi = coef.im * sqrt_power;
        ar = adj.g_inc[slot_id] * 2.0f * fr + adj.g_f_inc * fr;
        ai = adj.g_inc[slot_id] * 2.0f * fi + adj.g_f_inc * fi;
    }
}

// d(contribution) along one plain (value-only) field derivative of a
// diffraction path: receiver projection plus the diffraction-vector adjoint
__device__ __forceinline__ float contract_plain(const C3<float>& dfield, const V3<float>& arrival,
                                                const V3<float>& rx_pol, float sqrt_power, float ar, float ai,
                                                const CellAdjoint& adj) {
    const Cx<float> dc = project_receiver(dfield, arrival, rx_pol);
    float v = ar * dc.re + ai * dc.im;
    const Cx<float> comps[3] = {dfield.x, dfield.y, dfield.z};
#pragma unroll
    for (int j = 0; j < 3; ++j) v += adj.b[j].re * comps[j].re + adj.b[j].im * comps[j].im;
    return v * sqrt_power;
}

// Diffraction adjoint: forward-mode tangents over (tx xyz, frequency) only;
// the four face-material parameters enter through the face reflection
// coefficients, in which the field is linear, so their derivatives are the
// field's coefficient partials times d r / d(eps_r, sigma) of a two-tangent
// slab evaluation.
__device__ void diffraction_adjoint(const SceneView& s, const Edge& e, const PathParams& pp, const V3<float>& tx,
                                    const V3<float>& rx, const V3<float>& tx_pol, const V3<float>& rx_pol,
                                    float freq, double k_d, double length_d, float sqrt_power, int slot_id,
                                    const CellAdjoint& adj, LaneGrad& g) {
    using D4 = Dual<4>;
    using D2 = Dual<2>;
    const V3<D4> txd{seed<4>(tx.x, 0), seed<4>(tx.y, 1), seed<4>(tx.z, 2)};
    const D4 fd = seed<4>(freq, 3);
    const D4 eps[2] = {lift<D4>(pp.eps[0]), lift<D4>(pp.eps[1])};
    const D4 sig[2] = {lift<D4>(pp.sig[0]), lift<D4>(pp.sig[1])};
    V3<D4> arrival;
    DiffFaceTerms faces;
    const C3<D4> field = diffraction_tangent(s, e, txd, rx, tx_pol, fd, eps, sig, k_d, length_d, arrival, &faces);
    const Cx<D4> coef = project_receiver(field, arrival, lift3<D4>(rx_pol));
    float ar, ai;
    coef_weights(cval(coef), sqrt_power, slot_id, adj, ar, ai);
    float d[8];
    const Cx<D4> comps[3] = {field.x, field.y, field.z};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = ar * coef.re.d[i] + ai * coef.im.d[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) v += adj.b[j].re * comps[j].re.d[i] + adj.b[j].im * comps[j].im.d[i];
        d[i] = v * sqrt_power;
    }
    const V3<float> arr = vval(arrival);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int m = pp.material[k];
        Cx<D2> rs, rh;
        slab_fresnel<D2>(lift<D2>(faces.cos_face[k]), seed<2>(pp.eps[k], 0), seed<2>(pp.sig[k], 1),
                         lift<D2>(s.mat_mu[m]), lift<D2>(s.mat_gain[m]), lift<D2>(s.mat_thickness[m]),
                         lift<D2>(freq), rs, rh);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const C3<float> df = c3_scale(faces.d_r[2 * k], Cx<float>{rs.re.d[j], rs.im.d[j]}) +
                                 c3_scale(faces.d_r[2 * k + 1], Cx<float>{rh.re.d[j], rh.im.d[j]});
            d[4 + 2 * k + j] = contract_plain(df, arr, rx_pol, sqrt_power, ar, ai, adj);
        }
    }
    g.tx[0] += d[0];
    g.tx[1] += d[1];
    g.tx[2] += d[2];
    g.freq += d[3];
    add_material(g, pp.material[0], d[4], d[5]);
    add_material(g, pp.material[1], d[6], d[7]);
}

__device__ __forceinline__ void scatter(LaneGrad& g, const PathParams& pp, int pass, const float d[kTangents]) {
#pragma unroll
    for (int i = 0; i < kTangents; ++i) {
        const int slot = pass * kTangents + i;
        if (slot >= pp.count) break;
        if (slot < 3) {
            g.tx[slot] += d[i];
        } else if (slot == 3) {
            g.freq += d[i];
        } else {
            const int k = (slot - 4) >> 1;
            if ((slot - 4) & 1) add_material(g, pp.material[k], 0.0f, d[i]);
            else add_material(g, pp.material[k], d[i], 0.0f);
        }
    }
}

enum VjpPhase : int { kPhaseMain = 0, kPhaseDiff = 1 };

// Everything one stored path (descriptor `desc`) of `cell` needs for its
// adjoint: the cell adjoint, the endpoints, the frozen interactions (chain
// groups and the faces they were hit on, or the walls of a transmission, or
// the edge) and the local parameter slots.
struct PathSetup {
    CellAdjoint adj;
    V3<float> tx, tx_pol, rx, rx_pol;
    float sqrt_power, freq;
    double k_d, length_d;
    int comp, cand, depth, slot_id;
    int groups[kMaxDepth];
    int faces[kMaxDepth];
    Hit walls[kMaxHits];
    PathParams pp;
};

template <int kPhase>
__device__ __forceinline__ void setup_path(const SceneView& s, const wc_solve& q, const wc_adjoint& ga, int t,
                                           int cell, int desc, PathSetup& ps) {
    ps.adj = cell_adjoint(ga, q, cell);
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    ps.tx = tx;
    ps.tx_pol = ld3(q.tx_pol + 3 * t);
    ps.sqrt_power = sqrtf(fmaxf(q.tx_power[t], 0.0f));
    const int rx_index = q.rx_shared ? cell - q.cell_offset[t] : cell;
    const V3<float> rx = ld3(q.rx_pos + 3 * rx_index);
    ps.rx = rx;
    ps.rx_pol = ld3(q.rx_pol + 3 * (q.rx_pol_stride ? rx_index : 0));
    ps.freq = (float)q.frequency_hz;
    ps.k_d = q.frequency_hz * kTwoPi / kC0;
    const int comp = desc >> 24;
    ps.comp = comp;
    int cand = desc & 0xFFFFFF;
    if (cand == 0xFFFFFF) cand = -1;
    ps.cand = cand;
    PathParams& pp = ps.pp;
    pp.count = 4;
    int depth = 0;
    double length_d = 0.0;
    if constexpr (kPhase == kPhaseDiff) {
        const Edge e = s.edges[cand];
        DiffGeometry dg;
        diffraction_geometry(e, tx, rx, dg);
        length_d = dg.s + dg.sp;
        const int mn = e.material1 >= 0 ? e.material1 : e.material0;
        pp.material[0] = e.material0;
        pp.material[1] = mn;
        pp.eps[0] = s.mat_eps[e.material0];
        pp.sig[0] = s.mat_sigma[e.material0];
        pp.eps[1] = s.mat_eps[mn];
        pp.sig[1] = s.mat_sigma[mn];
        pp.count = 8;
    } else if (comp == kCidRefl) {
        const int* cg = q.chain_groups + cand * kMaxDepth;
        const double* img = q.chain_images + cand * kMaxDepth * 3;
        while (depth < kMaxDepth && cg[depth] >= 0) {
            ps.groups[depth] = cg[depth];
            ++depth;
        }
        // interaction faces of the frozen chain (image back-projection only)
        V3<double> next = to_d(rx);
        for (int k = depth - 1; k >= 0; --k) {
            double off;
            const V3<double> n = group_normal(s, ps.groups[k], off);
            const V3<double> im{img[3 * k], img[3 * k + 1], img[3 * k + 2]};
            const double si = dot(n, im) - off, sn = dot(n, next) - off;
            const V3<double> h = im + (next - im) * (si / (si - sn));
            int f = face_containing(s, ps.groups[k], to_f(h), to_f(n));
            ps.faces[k] = f >= 0 ? f : s.group_faces[s.group_face_offset[ps.groups[k]]];
            next = h;
        }
        const V3<double> last{img[3 * (depth - 1)], img[3 * (depth - 1) + 1], img[3 * (depth - 1) + 2]};
        length_d = norm(to_d(rx) - last);
        for (int k = 0; k < depth; ++k) {
            const int m = s.face_material[ps.faces[k]];
            pp.material[k] = m;
            pp.eps[k] = s.mat_eps[m];
            pp.sig[k] = s.mat_sigma[m];
        }
        pp.count = 4 + 2 * depth;
    } else if (comp == kCidTrans) {
        const V3<float> seg = rx - tx;
        const float len = norm(seg);
        int wall_count = 0;
        if (len > 2.0f * kSegmentBias) {
            const float eps = kSegmentBias / len;
            const RayPre r = ray_setup(tx.x, tx.y, tx.z, seg.x, seg.y, seg.z);
            wall_count = gather_hits(s, r, eps, 1.0f - eps, q.max_depth, ps.walls);
        }
        depth = min(wall_count, q.max_depth);
        double carrier = norm(to_d(rx) - to_d(tx));
        for (int w = 0; w < depth; ++w) {
            double off;
            const V3<double> nd = group_normal(s, ps.walls[w].group, off);
            const V3<double> dir = (to_d(rx) - to_d(tx)) * (1.0 / norm(to_d(rx) - to_d(tx)));
            const double ct = fmin(fmax(fabs(dot(dir, nd)), 1e-6), 1.0);
            const int m = s.face_material[ps.walls[w].face];
            double wall = 0.0;
            for (int l = 0; l < s.mat_layer_count[m]; ++l)
                wall += fmax((double)s.layer_thickness[s.mat_layer_offset[m] + l], 0.0);
            carrier -= wall * ct;
            pp.material[w] = m;
            pp.eps[w] = s.layer_eps[s.mat_layer_offset[m]];
            pp.sig[w] = s.layer_sigma[s.mat_layer_offset[m]];
        }
        length_d = carrier;
        pp.count = 4 + 2 * depth;
    } else {
        length_d = norm(to_d(rx) - to_d(tx));
    }
    ps.depth = depth;
    ps.length_d = length_d;
    ps.slot_id = slot_of_component(comp);
}

// Adjoint of one stored path (descriptor `desc`) of `cell`, accumulated into
// g. kPhaseMain handles line of sight, reflection chains and transmission,
// kPhaseDiff diffraction: each phase is its own kernel with its own register
// budget.
template <int kPhase>
__device__ void adjoint_path(const SceneView& s, const wc_solve& q, const wc_adjoint& ga, int t, int cell, int desc,
                             LaneGrad& g) {
    PathSetup ps;
    setup_path<kPhase>(s, q, ga, t, cell, desc, ps);
    const PathParams& pp = ps.pp;
    if constexpr (kPhase == kPhaseDiff) {
        diffraction_adjoint(s, s.edges[ps.cand], pp, ps.tx, ps.rx, ps.tx_pol, ps.rx_pol, ps.freq, ps.k_d,
                            ps.length_d, ps.sqrt_power, ps.slot_id, ps.adj, g);
        return;
    }
    const int comp = ps.comp;
    const int passes = (pp.count + kTangents - 1) / kTangents;
    for (int pass = 0; pass < passes; ++pass) {
        const V3<D8> txd{seeded(ps.tx.x, 0, pass), seeded(ps.tx.y, 1, pass), seeded(ps.tx.z, 2, pass)};
        const D8 fd = seeded(ps.freq, 3, pass);
        D8 eps[kMaxDepth], sig[kMaxDepth];
        const int nm = (pp.count - 4) / 2;
        for (int k = 0; k < nm; ++k) {
            eps[k] = seeded(pp.eps[k], 4 + 2 * k, pass);
            sig[k] = seeded(pp.sig[k], 5 + 2 * k, pass);
        }
        V3<D8> arrival;
        C3<D8> field;
        {
            if (comp == kCidLos) {
                field = los_tangent(txd, ps.rx, ps.tx_pol, fd, ps.k_d, ps.length_d, arrival);
            } else if (comp == kCidRefl) {
                field = reflection_tangent(s, ps.depth, ps.groups, ps.faces, txd, ps.rx, ps.tx_pol, fd, eps, sig,
                                           ps.k_d, ps.length_d, arrival);
            } else {
                field = transmission_tangent(s, ps.depth, ps.walls, txd, ps.rx, ps.tx_pol, fd, eps, sig, ps.k_d,
                                             ps.length_d, arrival);
            }
        }
        float d[kTangents];
        contract(field, arrival, ps.rx_pol, ps.sqrt_power, ps.slot_id, ps.adj, false, d);
        scatter(g, pp, pass, d);
    }
}

// Geometry adjoint outputs (float64 accumulators, caller-zeroed):
// dL/d(group plane n xyz, c) [G, 4] and dL/d(edge origin xyz, frame rotation
// xyz, length, wedge index) [E, 8].
struct GeoGrad {
    double* plane;
    double* edge;
};

constexpr int kGeoTangents = 8;
using DG = Dual<kGeoTangents>;

__device__ __forceinline__ DG geo_seed(float v, int slot, int pass) {
    const int local = slot - pass * kGeoTangents;
    return seed<kGeoTangents>(v, (local >= 0 && local < kGeoTangents) ? local : -1);
}

// Geometry adjoint of one stored path: the same recomputation with the
// interaction geometry carrying the tangents (4 per reflecting or transmitted
// plane, 8 per edge); the contracted derivatives are added to the per-group
// and per-edge accumulators with float64 atomics. Line-of-sight paths have no
// geometry.
template <int kPhase>
__device__ void geometry_path(const SceneView& s, const wc_solve& q, const wc_adjoint& ga, int t, int cell, int desc,
                              const GeoGrad& out) {
    PathSetup ps;
    setup_path<kPhase>(s, q, ga, t, cell, desc, ps);
    float d[kGeoTangents];
    if constexpr (kPhase == kPhaseDiff) {
        const Edge e = s.edges[ps.cand];
        EdgeGeom<DG> g;
        g.A = {seed<kGeoTangents>(e.ax, 0), seed<kGeoTangents>(e.ay, 1), seed<kGeoTangents>(e.az, 2)};
        // an infinitesimal rotation w of the frame: v -> v + w x v
        const V3<DG> w{seed<kGeoTangents>(0.0f, 3), seed<kGeoTangents>(0.0f, 4), seed<kGeoTangents>(0.0f, 5)};
        const V3<DG> d0 = v3c<DG>(e.dx, e.dy, e.dz), t0 = v3c<DG>(e.t0x, e.t0y, e.t0z);
        const V3<DG> n0 = v3c<DG>(e.n0x, e.n0y, e.n0z);
        g.d = d0 + cross(w, d0);
        g.t0 = t0 + cross(w, t0);
        g.n0 = n0 + cross(w, n0);
        g.length = seed<kGeoTangents>(e.length, 6);
        g.n = seed<kGeoTangents>(e.n, 7);
        V3<DG> arrival;
        const C3<DG> field = diffraction_tangent_edge(s, e, g, ps.tx, ps.rx, ps.tx_pol, ps.freq, ps.k_d,
                                                      ps.length_d, arrival);
        contract(field, arrival, ps.rx_pol, ps.sqrt_power, ps.slot_id, ps.adj, true, d);
        double* dst = out.edge + (size_t)kGeoTangents * ps.cand;
#pragma unroll
        for (int i = 0; i < kGeoTangents; ++i)
            if (d[i] != 0.0f) unsafeAtomicAdd(dst + i, (double)d[i]);
        return;
    }
    if (ps.comp == kCidLos || ps.depth == 0) return;
    const bool refl = ps.comp == kCidRefl;
    const int slots = 4 * ps.depth;
    for (int pass = 0; pass * kGeoTangents < slots; ++pass) {
        V3<DG> nrm[kMaxDepth];
        DG off[kMaxDepth];
        for (int k = 0; k < ps.depth; ++k) {
            double o;
            const V3<float> n = to_f(group_normal(s, refl ? ps.groups[k] : ps.walls[k].group, o));
            nrm[k] = {geo_seed(n.x, 4 * k, pass), geo_seed(n.y, 4 * k + 1, pass), geo_seed(n.z, 4 * k + 2, pass)};
            off[k] = geo_seed((float)o, 4 * k + 3, pass);
        }
        V3<DG> arrival;
        const C3<DG> field =
            refl ? reflection_tangent_planes(s, ps.depth, ps.faces, nrm, off, ps.tx, ps.rx, ps.tx_pol, ps.freq,
                                             ps.k_d, ps.length_d, arrival)
                 : transmission_tangent_walls(s, ps.depth, ps.walls, nrm, ps.tx, ps.rx, ps.tx_pol, ps.freq, ps.k_d,
                                              ps.length_d, arrival);
        contract(field, arrival, ps.rx_pol, ps.sqrt_power, ps.slot_id, ps.adj, false, d);
#pragma unroll
        for (int i = 0; i < kGeoTangents; ++i) {
            const int slot = pass * kGeoTangents + i;
            if (slot >= slots || d[i] == 0.0f) continue;
            const int k = slot >> 2;
            const int group = refl ? ps.groups[k] : ps.walls[k].group;
            unsafeAtomicAdd(out.plane + 4 * (size_t)group + (slot & 3), (double)d[i]);
        }
    }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ int wave_inclusive_scan(int v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int n = __shfl_up(v, off);
        if (lane >= off) v += n;
    }
    return v;
}

// Occupancy targets (waves per SIMD) of the adjoint phases. The line-of-sight /
// reflection / transmission phase runs 2 waves (256 VGPRs and more private
// memory, but its many short paths hide the latency: scaled-city step
// 1434 -> 1372 ms, canyon 3.26 -> 3.30 ms); the diffraction phase and the
// geometry adjoint keep the compiler's choice (1: the diffraction adjoint at 2
// waves took the canyon's 1.13 ms to 1.43 ms).
#ifndef WC_VJP_WAVES_PER_EU
#define WC_VJP_WAVES_PER_EU 1
#endif
#ifndef WC_VJP_MAIN_WAVES_PER_EU
#define WC_VJP_MAIN_WAVES_PER_EU 2
#endif

// Work items of the adjoint per tile: the tile's flattened path list is cut
// into this many slices of (nearly) equal iteration counts, one wavefront
// each, so a tile holding many paths (cells near the transmitter see hundreds
// of edges) spreads over many SIMDs while the slices of light tiles end at
// once. Each (tile, slice) owns one partial row.
#ifndef WC_VJP_SLICES
#define WC_VJP_SLICES 16
#endif
constexpr int kVjpSlices = WC_VJP_SLICES;
// The line-of-sight / reflection / transmission phase deals each tile's paths
// over fewer work items: its tiles hold tens to hundreds of paths, not the
// thousands of diffraction paths near the transmitter, and a work item's fixed
// cost (descriptor scan, 2M + 4 wavefront sums, at one wavefront per SIMD)
// then dominates. Its work items fill the first of the tile's partial rows
// and zero the rest, which the diffraction phase adds to. The geometry adjoint
// of the same paths (no partial rows) is dealt the same way.
#ifndef WC_VJP_MAIN_SLICES
#define WC_VJP_MAIN_SLICES 4
#endif
constexpr int kVjpMainSlices = WC_VJP_MAIN_SLICES;
static_assert(kVjpMainSlices >= 1 && kVjpMainSlices <= kVjpSlices, "main-phase slices");

// One wavefront per (tile of 64 units, slice) (unit = one lane-segment of a
// cell's descriptors; 64 / L cells). The tile's paths of this phase are
// flattened over its units (wave prefix sum of their counts) and dealt to the
// 64 lanes round-robin; slice k takes iterations [k I / S, (k + 1) I / S) of
// the I = ceil(paths / 64) the tile needs. Within a unit the reflection pass
// stored its descriptors before the diffraction pass appended its own.
// kPhaseMain writes the item's partial vector, kPhaseDiff adds to it. With
// kGeometry the same dealing drives geometry_path instead (no partials).
template <int kPhase, bool kGeometry = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(kPhase == 0 && !kGeometry ? WC_VJP_MAIN_WAVES_PER_EU : WC_VJP_WAVES_PER_EU)))
vjp_kernel(SceneView s, wc_solve q, wc_adjoint ga, const int* __restrict__ tiles, int tile_count, GeoGrad geo,
           int slices) {
    const int lane = threadIdx.x & 63;
    const int work = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (work >= tile_count * slices) return;
    const int tile = work / slices;
    const int slice = work - tile * slices;
    const int item = tile * kVjpSlices + slice;  // the work item's partial row
    const int width = q.lanes_per_cell;
    const int t = tiles[2 * tile];
    const int pos = tiles[2 * tile + 1] + lane / width;
    const int cap_sub = ga.topology_capacity / width;
    int cell = 0, n_main = 0, n_total = 0;
    const int* seg = ga.topology;
    if (pos < q.cell_offset[t + 1]) {
        cell = q.cell_order ? q.cell_order[pos] : pos;
        seg = ga.topology + (size_t)cell * ga.topology_capacity + (size_t)(lane & (width - 1)) * cap_sub;
        if (width == 1) {
            // one segment per cell: its length is the path count (capped), and
            // the diffraction descriptors, appended last, start at the first
            // index whose component is diffraction (binary search)
            n_total = min(ga.path_count[cell], cap_sub);
            int lo = 0, hi = n_total;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((seg[mid] >> 24) == kCidDiff) hi = mid;
                else lo = mid + 1;
            }
            n_main = lo;
        } else {
            int first_diff = -1;
            while (n_total < cap_sub) {
                const int d = seg[n_total];
                if (d == -1) break;
                if (first_diff < 0 && (d >> 24) == kCidDiff) first_diff = n_total;
                ++n_total;
            }
            n_main = first_diff < 0 ? n_total : first_diff;
        }
    }
    const int mine = kPhase == kPhaseMain ? n_main : n_total - n_main;
    const int first = kPhase == kPhaseMain ? 0 : n_main;
    const int inclusive = wave_inclusive_scan(mine, lane);
    const int exclusive = inclusive - mine;
    const int total = __shfl(inclusive, 63);
    LaneGrad g;
#pragma unroll
    for (int i = 0; i < kMaxMaterials; ++i) {
        g.eps[i] = 0.0f;
        g.sig[i] = 0.0f;
    }
    g.freq = 0.0f;
    g.tx[0] = g.tx[1] = g.tx[2] = 0.0f;
    const int iterations = (total + 63) >> 6;
    const int it0 = (int)((long long)iterations * slice / slices);
    const int it1 = (int)((long long)iterations * (slice + 1) / slices);
    for (int base = it0 * 64; base < it1 * 64; base += 64) {
        // owner unit u of path j: the last unit whose exclusive prefix is <= j
        // (all lanes take part in the cross-lane reads)
        const int j = base + lane;
        int u = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const int probe = u + step;
            if (__shfl(exclusive, probe) <= j) u = probe;
        }
        const int u_cell = __shfl(cell, u);
        const int u_first = __shfl(first, u);
        const int u_excl = __shfl(exclusive, u);
        const int u_sub = u & (width - 1);
        if (j < total) {
            const int* useg = ga.topology + (size_t)u_cell * ga.topology_capacity + (size_t)u_sub * cap_sub;
            if constexpr (kGeometry) geometry_path<kPhase>(s, q, ga, t, u_cell, useg[u_first + (j - u_excl)], geo);
            else adjoint_path<kPhase>(s, q, ga, t, u_cell, useg[u_first + (j - u_excl)], g);
        }
    }
    if constexpr (kGeometry) return;
    const int m = s.material_count;
    const int stride = 2 * m + 4;
    float* out = ga.partials + (size_t)item * stride;
    if (kPhase == kPhaseMain && slice == 0) {
        // the tile's rows this phase's work items do not cover
        for (int r = slices; r < kVjpSlices; ++r)
            for (int j = lane; j < stride; j += 64) ga.partials[((size_t)tile * kVjpSlices + r) * stride + j] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < kMaxMaterials; ++i) {
        const float se = wave_sum(g.eps[i]);
        const float ss = wave_sum(g.sig[i]);
        if (lane == 0 && i < m) {
            if (kPhase == kPhaseMain) {
                out[i] = se;
                out[m + i] = ss;
            } else {
                out[i] += se;
                out[m + i] += ss;
            }
        }
    }
    const float sf = wave_sum(g.freq);
    const float sx = wave_sum(g.tx[0]), sy = wave_sum(g.tx[1]), sz = wave_sum(g.tx[2]);
    if (lane == 0) {
        if (kPhase == kPhaseMain) {
            out[2 * m] = sf;
            out[2 * m + 1] = sx;
            out[2 * m + 2] = sy;
            out[2 * m + 3] = sz;
        } else {
            out[2 * m] += sf;
            out[2 * m + 1] += sx;
            out[2 * m + 2] += sy;
            out[2 * m + 3] += sz;
        }
    }
}

// Column sums of the partial rows (kVjpSlices per tile): one workgroup per
// column. Thread j sums rows j, j + 256, ... in float64, then a fixed-shape
// tree in LDS combines the 256 partial sums, so the result is deterministic.
// Material and frequency columns sum every row; transmitter columns only the
// rows of their transmitter's tiles.
__global__ void __launch_bounds__(256) vjp_reduce_kernel(const float* __restrict__ partials, int tile_count,
                                                         int stride, const int* __restrict__ tiles, int tx_count,
                                                         float* __restrict__ grad_params,
                                                         float* __restrict__ grad_tx) {
    __shared__ double sums[256];
    const int m2 = stride - 3;  // 2M + 1 shared columns
    const int column = blockIdx.x;
    const int tx = column < m2 ? -1 : (column - m2) / 3;
    const int src = column < m2 ? column : m2 + (column - m2) % 3;
    double acc = 0.0;
    const int rows = tile_count * kVjpSlices;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        if (tx < 0 || tiles[2 * (r / kVjpSlices)] == tx) acc += (double)partials[(size_t)r * stride + src];
    }
    sums[threadIdx.x] = acc;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half) sums[threadIdx.x] += sums[threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (column < m2) grad_params[column] = (float)sums[0];
        else grad_tx[column - m2] = (float)sums[0];
    }
}

// ------------------------------------------------------------------ JVP
// Tangent of the radiomap along one direction of its parameters (per
// material first-layer eps_r and sigma, the carrier frequency, every
// transmitter position): each stored path is recomputed with one-tangent dual
// numbers seeded by the direction, and its path-field tangent is summed per
// cell and slot in float64, then mapped through the forward's finalize
// (coherent |S|^2 or incoherent sums).
using D1 = Dual<1>;

__device__ __forceinline__ D1 jv(float v, float t) {
    D1 r = make_dual<1>(v);
    r.d[0] = t;
    return r;
}

struct JvpDirection {
    const float* d_eps;   // [M] or null
    const float* d_sig;   // [M] or null
    float d_freq;
    const float* d_tx;    // [T, 3] or null
};

// path field value and tangent (times sqrt power), its slot and, for a
// diffraction path, the field 3-vector value and tangent (times sqrt power)
__device__ void path_jvp(const SceneView& s, const wc_solve& q, const wc_adjoint& ga, int t, int cell, int desc,
                         const JvpDirection& dir, Cx<float>& pf, Cx<float>& dpf, C3<float>& vec, C3<float>& dvec,
                         int& slot) {
    const int comp = desc >> 24;
    PathSetup ps;
    if (comp == kCidDiff) setup_path<kPhaseDiff>(s, q, ga, t, cell, desc, ps);
    else setup_path<kPhaseMain>(s, q, ga, t, cell, desc, ps);
    const PathParams& pp = ps.pp;
    const V3<float> dt = dir.d_tx ? ld3(dir.d_tx + 3 * t) : v3c<float>(0.f, 0.f, 0.f);
    const V3<D1> txd{jv(ps.tx.x, dt.x), jv(ps.tx.y, dt.y), jv(ps.tx.z, dt.z)};
    const D1 fd = jv(ps.freq, dir.d_freq);
    auto d_of = [&](const float* a, int m) { return (a && m >= 0) ? a[m] : 0.0f; };
    V3<D1> arrival;
    C3<D1> field;
    if (comp == kCidDiff) {
        const Edge e = s.edges[ps.cand];
        const D1 eps[2] = {lift<D1>(pp.eps[0]), lift<D1>(pp.eps[1])};
        const D1 sig[2] = {lift<D1>(pp.sig[0]), lift<D1>(pp.sig[1])};
        DiffFaceTerms faces;
        field = diffraction_tangent(s, e, txd, ps.rx, ps.tx_pol, fd, eps, sig, ps.k_d, ps.length_d, arrival, &faces);
        // the face materials enter through the faces' slab coefficients
        for (int k = 0; k < 2; ++k) {
            const int m = pp.material[k];
            Cx<Dual<2>> rs, rh;
            slab_fresnel<Dual<2>>(lift<Dual<2>>(faces.cos_face[k]), seed<2>(pp.eps[k], 0), seed<2>(pp.sig[k], 1),
                                  lift<Dual<2>>(s.mat_mu[m]), lift<Dual<2>>(s.mat_gain[m]),
                                  lift<Dual<2>>(s.mat_thickness[m]), lift<Dual<2>>(ps.freq), rs, rh);
            const float de = d_of(dir.d_eps, m), ds = d_of(dir.d_sig, m);
            const Cx<float> drs{rs.re.d[0] * de + rs.re.d[1] * ds, rs.im.d[0] * de + rs.im.d[1] * ds};
            const Cx<float> drh{rh.re.d[0] * de + rh.re.d[1] * ds, rh.im.d[0] * de + rh.im.d[1] * ds};
            const C3<float> df = c3_scale(faces.d_r[2 * k], drs) + c3_scale(faces.d_r[2 * k + 1], drh);
            field.x.re.d[0] += df.x.re; field.x.im.d[0] += df.x.im;
            field.y.re.d[0] += df.y.re; field.y.im.d[0] += df.y.im;
            field.z.re.d[0] += df.z.re; field.z.im.d[0] += df.z.im;
        }
    } else {
        D1 eps[kMaxDepth], sig[kMaxDepth];
        const int nm = (pp.count - 4) / 2;
        for (int k = 0; k < nm; ++k) {
            eps[k] = jv(pp.eps[k], d_of(dir.d_eps, pp.material[k]));
            sig[k] = jv(pp.sig[k], d_of(dir.d_sig, pp.material[k]));
        }
        if (comp == kCidLos) field = los_tangent(txd, ps.rx, ps.tx_pol, fd, ps.k_d, ps.length_d, arrival);
        else if (comp == kCidRefl)
            field = reflection_tangent(s, ps.depth, ps.groups, ps.faces, txd, ps.rx, ps.tx_pol, fd, eps, sig, ps.k_d,
                                       ps.length_d, arrival);
        else
            field = transmission_tangent(s, ps.depth, ps.walls, txd, ps.rx, ps.tx_pol, fd, eps, sig, ps.k_d,
                                         ps.length_d, arrival);
    }
    const Cx<D1> coef = project_receiver(field, arrival, lift3<D1>(ps.rx_pol));
    const float sp = ps.sqrt_power;
    pf = {coef.re.v * sp, coef.im.v * sp};
    dpf = {coef.re.d[0] * sp, coef.im.d[0] * sp};
    vec = {{field.x.re.v * sp, field.x.im.v * sp}, {field.y.re.v * sp, field.y.im.v * sp},
           {field.z.re.v * sp, field.z.im.v * sp}};
    dvec = {{field.x.re.d[0] * sp, field.x.im.d[0] * sp}, {field.y.re.d[0] * sp, field.y.im.d[0] * sp},
            {field.z.re.d[0] * sp, field.z.im.d[0] * sp}};
    slot = ps.slot_id;
}

// one thread per cell: its stored descriptors in order
__global__ void __launch_bounds__(256) jvp_kernel(SceneView s, wc_solve q, wc_adjoint ga, JvpDirection dir,
                                                  float* __restrict__ d_power, float* __restrict__ d_field,
                                                  float* __restrict__ d_cpower, float* __restrict__ d_cfield,
                                                  float* __restrict__ d_dpower) {
    const int cells = q.total_cells;
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x; cell < cells; cell += blockDim.x * gridDim.x) {
        int lo = 0, hi = q.tx_count - 1;  // transmitter: last t with cell_offset[t] <= cell
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (q.cell_offset[mid] <= cell) lo = mid;
            else hi = mid - 1;
        }
        const int t = lo;
        double ds_re[4] = {0, 0, 0, 0}, ds_im[4] = {0, 0, 0, 0}, dp_inc[4] = {0, 0, 0, 0};
        double dv[6] = {0, 0, 0, 0, 0, 0};
        const int width = q.lanes_per_cell;
        const int cap_sub = ga.topology_capacity / width;
        int remaining = ga.path_count[cell];
        for (int sub = 0; sub < width; ++sub) {
            const int* seg = ga.topology + (size_t)cell * ga.topology_capacity + (size_t)sub * cap_sub;
            const int n = width == 1 ? min(remaining, cap_sub) : cap_sub;
            for (int j = 0; j < n; ++j) {
                const int desc = seg[j];
                if (desc == -1) break;
                Cx<float> pf, dpf;
                C3<float> vec, dvec;
                int sl;
                path_jvp(s, q, ga, t, cell, desc, dir, pf, dpf, vec, dvec, sl);
                if (sl < 0 || sl > 3) continue;
                ds_re[sl] += dpf.re;
                ds_im[sl] += dpf.im;
                dp_inc[sl] += 2.0 * ((double)pf.re * dpf.re + (double)pf.im * dpf.im);
                if ((desc >> 24) == kCidDiff) {
                    dv[0] += dvec.x.re; dv[1] += dvec.x.im; dv[2] += dvec.y.re;
                    dv[3] += dvec.y.im; dv[4] += dvec.z.re; dv[5] += dvec.z.im;
                }
            }
        }
        double s_re = 0.0, s_im = 0.0, p_inc = 0.0;
        for (int sl = 0; sl < 4; ++sl) {
            const size_t o = (size_t)sl * cells + cell;
            const double c_re = ga.component_field[2 * o], c_im = ga.component_field[2 * o + 1];
            d_cfield[2 * o] = (float)ds_re[sl];
            d_cfield[2 * o + 1] = (float)ds_im[sl];
            d_cpower[o] = (float)(q.coherent ? 2.0 * (c_re * ds_re[sl] + c_im * ds_im[sl]) : dp_inc[sl]);
            s_re += ds_re[sl];
            s_im += ds_im[sl];
            p_inc += dp_inc[sl];
        }
        if (q.coherent) {
            const double f_re = ga.field_total[2 * cell], f_im = ga.field_total[2 * cell + 1];
            d_power[cell] = (float)(2.0 * (f_re * s_re + f_im * s_im));
            d_field[2 * cell] = (float)s_re;
            d_field[2 * cell + 1] = (float)s_im;
        } else {
            const double p = ga.power_total ? ga.power_total[cell] : 0.0;
            d_power[cell] = (float)p_inc;
            d_field[2 * cell] = p > 0.0 ? (float)(p_inc / (2.0 * sqrt(p))) : 0.0f;
            d_field[2 * cell + 1] = 0.0f;
        }
        if (d_dpower) {
            double acc = 0.0;
            if (ga.diffraction_vector)
                for (int j = 0; j < 6; ++j) acc += 2.0 * (double)ga.diffraction_vector[6 * cell + j] * dv[j];
            d_dpower[cell] = (float)acc;
        }
    }
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_vjp_reduce(const float* partials, int tile_count, int stride, const int* tiles, int tx_count,
                             float* grad_params, float* grad_tx, void* stream) {
    if (stride < 4 || tile_count < 0 || tx_count < 0) return wc_fail("wc_vjp_reduce: invalid sizes");
    if (!partials || !tiles || !grad_params || !grad_tx) return wc_fail("wc_vjp_reduce: null argument");
    const int columns = (stride - 3) + 3 * tx_count;
    hipLaunchKernelGGL(vjp_reduce_kernel, dim3(columns), dim3(256), 0, (hipStream_t)stream, partials, tile_count,
                       stride, tiles, tx_count, grad_params, grad_tx);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_vjp_reduce", err);
}

extern "C" int wc_vjp_partial_stride(int material_count) { return 2 * material_count + 4; }

extern "C" int wc_vjp_partial_rows(int tile_count) { return tile_count * kVjpSlices; }

static int vjp_launch(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint, const int* tiles,
                      int tile_count, int phases, const char* who, void* stream, const GeoGrad* geo = nullptr) {
    if (!scene || !solve || !adjoint) return wc_fail("%s: null argument", who);
    if (scene->material_count > kMaxMaterials)
        return wc_fail("%s: at most 16 materials are supported by the adjoint", who);
    if (!adjoint->topology || !adjoint->path_count || adjoint->topology_capacity <= 0)
        return wc_fail("%s: the forward topology (descriptors and counts) is required", who);
    if (!adjoint->field_total || !adjoint->component_field || (!geo && !adjoint->partials))
        return wc_fail("%s: forward fields and the partials buffer are required", who);
    if (!tiles || tile_count < 0) return wc_fail("%s: tiles are required", who);
    if (tile_count == 0) return 0;
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    if (adjoint->topology_capacity % solve->lanes_per_cell != 0)
        return wc_fail("%s: topology_capacity must be a multiple of lanes_per_cell", who);
    if (tile_count > 0x7fffffff / kVjpSlices) return wc_fail("%s: too many tiles (%d)", who, tile_count);
    const int blocks = (tile_count * kVjpSlices + 3) / 4;
    const int main_blocks = (tile_count * kVjpMainSlices + 3) / 4;
    const GeoGrad none{nullptr, nullptr};
    if (geo) {
        if (phases & 1)
            hipLaunchKernelGGL((vjp_kernel<kPhaseMain, true>), dim3(main_blocks), dim3(256), 0,
                               (hipStream_t)stream, s, *solve, *adjoint, tiles, tile_count, *geo, kVjpMainSlices);
        const hipError_t err = hipGetLastError();
        return err == hipSuccess ? 0 : wc_fail_hip(who, err);
    }
    if (phases & 1) {
        // a shared receiver grid concentrates its reflection paths next to the
        // transmitter (8 items per tile: canyon step 3.16 -> 3.08 ms); point
        // sets spread them (4: 64-Tx batch 3.42 vs 3.62 ms with 8)
        const int main_slices = solve->rx_shared ? 2 * kVjpMainSlices : kVjpMainSlices;
        hipLaunchKernelGGL(vjp_kernel<kPhaseMain>, dim3((tile_count * main_slices + 3) / 4), dim3(256), 0,
                           (hipStream_t)stream, s, *solve, *adjoint, tiles, tile_count, none,
                           main_slices <= kVjpSlices ? main_slices : kVjpSlices);
    }
    if ((phases & 2) && (solve->components & kCompDiff)) {
        // a shared receiver grid has the heavy tiles next to the transmitter
        // (16 work items each); per-transmitter point sets spread their
        // diffraction paths more evenly (8: 64-Tx batch step 3.43 -> 3.35 ms)
        const int diff_slices = solve->rx_shared ? kVjpSlices : (kVjpSlices >= 8 ? 8 : kVjpSlices);
        hipLaunchKernelGGL(vjp_kernel<kPhaseDiff>, dim3((tile_count * diff_slices + 3) / 4), dim3(256), 0,
                           (hipStream_t)stream, s, *solve, *adjoint, tiles, tile_count, none, diff_slices);
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_radiomap_vjp(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint,
                               const int* tiles, int tile_count, void* stream) {
    return vjp_launch(scene, solve, adjoint, tiles, tile_count, 3, "wc_radiomap_vjp", stream);
}

extern "C" int wc_radiomap_vjp_phase(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint,
                                     const int* tiles, int tile_count, int phase, void* stream) {
    if (phase != 1 && phase != 2) return wc_fail("wc_radiomap_vjp_phase: phase must be 1 or 2 (got %d)", phase);
    return vjp_launch(scene, solve, adjoint, tiles, tile_count, phase, "wc_radiomap_vjp_phase", stream);
}

extern "C" int wc_radiomap_vjp_geometry(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint,
                                        const int* tiles, int tile_count, double* grad_group_plane,
                                        double* grad_edge, void* stream) {
    if (!scene || !solve) return wc_fail("wc_radiomap_vjp_geometry: null argument");
    if (scene->group_count > 0 && !grad_group_plane)
        return wc_fail("wc_radiomap_vjp_geometry: grad_group_plane is required");
    // The edge part (geometry_path<kPhaseDiff>) is switched off: its kernel
    // stopped with a GPU memory aperture violation whose cause is not
    // identified (the same device code runs clean under AddressSanitizer in a
    // CPU build over every edge/receiver pair of that scene). Solves with
    // diffraction are refused rather than returning a partial gradient.
    if (scene->edge_count > 0 && (solve->components & kCompDiff))
        return wc_fail("wc_radiomap_vjp_geometry: vertex gradients of diffraction paths are not available; "
                       "solve without diffraction");
    (void)grad_edge;
    const GeoGrad geo{grad_group_plane, nullptr};
    return vjp_launch(scene, solve, adjoint, tiles, tile_count, 1, "wc_radiomap_vjp_geometry", stream, &geo);
}

namespace wc {
namespace {

template <int N> __device__ __forceinline__ V3<Dual<N>> seed3(const float* p, int first) {
    return {seed<N>(p[0], first), seed<N>(p[1], first + 1), seed<N>(p[2], first + 2)};
}

template <int N> __device__ __forceinline__ V3<Dual<N>> unit(const V3<Dual<N>>& v) {
    return v * (lift<Dual<N>>(1.0f) / norm(v));
}

// Transposed Jacobians of the scene geometry with respect to the vertices.
// Group g (plane of its representative face, canonical sign: largest normal
// component positive; c = n . v0) maps dL/d(n, c) to 9 vertex-coordinate
// contributions; edge e (vertices a, b, c0, c1) maps dL/d(A, rotation,
// length, n) to 12. Contributions are stored per (entity, vertex slot, xyz):
// groups first, [G, 3, 3], then edges, [E, 4, 3].
__global__ void __launch_bounds__(256) geometry_jacobian_kernel(
    const float* __restrict__ vertices, const int* __restrict__ faces, const int* __restrict__ group_face,
    int group_count, const double* __restrict__ grad_plane, const int* __restrict__ edge_vertices, int edge_count,
    const double* __restrict__ grad_edge, float* __restrict__ contrib) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < group_count) {
        using D = Dual<9>;
        const int f = group_face[i];
        const V3<D> v0 = seed3<9>(vertices + 3 * faces[3 * f], 0);
        const V3<D> v1 = seed3<9>(vertices + 3 * faces[3 * f + 1], 3);
        const V3<D> v2 = seed3<9>(vertices + 3 * faces[3 * f + 2], 6);
        V3<D> n = unit(cross(v1 - v0, v2 - v0));
        const float ax = fabsf(n.x.v), ay = fabsf(n.y.v), az = fabsf(n.z.v);
        const float major = (ax >= ay && ax >= az) ? n.x.v : (ay >= az ? n.y.v : n.z.v);
        if (major < 0.0f) n = -n;
        const D c = dot(n, v0);
        const double* g = grad_plane + 4 * (size_t)i;
        float* out = contrib + 9 * (size_t)i;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            out[k] = (float)(g[0] * n.x.d[k] + g[1] * n.y.d[k] + g[2] * n.z.d[k] + g[3] * c.d[k]);
        return;
    }
    const int e = i - group_count;
    if (e >= edge_count) return;
    using D = Dual<12>;
    const int* ev = edge_vertices + 4 * e;
    const V3<D> va = seed3<12>(vertices + 3 * ev[0], 0);
    const V3<D> vb = seed3<12>(vertices + 3 * ev[1], 3);
    const V3<D> vc = seed3<12>(vertices + 3 * ev[2], 6);
    const V3<D> dv = vb - va;
    const D length = norm(dv);
    const V3<D> d = dv * (lift<D>(1.0f) / length);
    auto in_plane = [&](const V3<D>& third) {
        const V3<D> w = third - va;
        return unit(w - d * dot(w, d));
    };
    const V3<D> t0 = in_plane(vc);
    const V3<D> n0raw = unit(cross(vb - va, vc - va));
    const V3<D> n0 = unit(n0raw - d * dot(n0raw, d));
    D n = lift<D>(2.0f);
    if (ev[3] >= 0) {
        const V3<D> t1 = in_plane(seed3<12>(vertices + 3 * ev[3], 9));
        // n = 2 - alpha / pi with alpha = atan2(-n0.t1, t0.t1) in [0, 2 pi): the
        // 2 pi wrap does not change the derivative
        n = lift<D>(2.0f) - w_atan2(-dot(n0, t1), dot(t0, t1)) * (float)(1.0 / kPi);
    }
    const double* g = grad_edge + 8 * (size_t)e;
    float* out = contrib + 9 * (size_t)group_count + 12 * (size_t)e;
    const V3<float> df{d.x.v, d.y.v, d.z.v}, tf{t0.x.v, t0.y.v, t0.z.v}, nf{n0.x.v, n0.y.v, n0.z.v};
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        // the frame moves by a rotation w = (d x dd + t0 x dt0 + n0 x dn0) / 2
        const V3<float> dd{d.x.d[k], d.y.d[k], d.z.d[k]};
        const V3<float> dt{t0.x.d[k], t0.y.d[k], t0.z.d[k]};
        const V3<float> dn{n0.x.d[k], n0.y.d[k], n0.z.d[k]};
        const V3<float> w = (cross(df, dd) + cross(tf, dt) + cross(nf, dn)) * 0.5f;
        const float dA[3] = {va.x.d[k], va.y.d[k], va.z.d[k]};
        out[k] = (float)(g[0] * dA[0] + g[1] * dA[1] + g[2] * dA[2] + g[3] * w.x + g[4] * w.y + g[5] * w.z +
                         g[6] * length.d[k] + g[7] * n.d[k]);
    }
}

// grad[v] = sum of the contribution triples listed for v (CSR, fixed order)
__global__ void __launch_bounds__(256) vertex_gather_kernel(const float* __restrict__ contrib,
                                                            const int* __restrict__ offset,
                                                            const int* __restrict__ slot, int vertex_count,
                                                            float* __restrict__ grad) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= vertex_count) return;
    double x = 0.0, y = 0.0, z = 0.0;
    for (int k = offset[v]; k < offset[v + 1]; ++k) {
        const float* c = contrib + 3 * (size_t)slot[k];
        x += c[0];
        y += c[1];
        z += c[2];
    }
    grad[3 * v] = (float)x;
    grad[3 * v + 1] = (float)y;
    grad[3 * v + 2] = (float)z;
}

}  // namespace
}  // namespace wc

extern "C" int wc_geometry_vertex_grad(const float* vertices, const int* faces, const int* group_face,
                                       int group_count, const double* grad_group_plane, const int* edge_vertices,
                                       int edge_count, const double* grad_edge, const int* csr_offset,
                                       const int* csr_slot, int vertex_count, float* contrib, float* grad_vertices,
                                       void* stream) {
    if (group_count < 0 || edge_count < 0 || vertex_count < 0)
        return wc_fail("wc_geometry_vertex_grad: invalid sizes");
    if (vertex_count == 0) return 0;
    if (!vertices || !csr_offset || !grad_vertices || ((group_count || edge_count) && (!contrib || !csr_slot)))
        return wc_fail("wc_geometry_vertex_grad: null argument");
    if (group_count && (!faces || !group_face || !grad_group_plane))
        return wc_fail("wc_geometry_vertex_grad: group tables are required");
    if (edge_count && (!edge_vertices || !grad_edge))
        return wc_fail("wc_geometry_vertex_grad: edge tables are required");
    const int entities = group_count + edge_count;
    if (entities > 0)
        hipLaunchKernelGGL(geometry_jacobian_kernel, dim3((entities + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                           vertices, faces, group_face, group_count, grad_group_plane, edge_vertices, edge_count,
                           grad_edge, contrib);
    hipLaunchKernelGGL(vertex_gather_kernel, dim3((vertex_count + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       contrib, csr_offset, csr_slot, vertex_count, grad_vertices);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_geometry_vertex_grad", err);
}

extern "C" int wc_radiomap_jvp(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* primal,
                               const wc_tangent_in* tangent, const wc_tangent_out* out, void* stream) {
    if (!scene || !solve || !primal || !tangent || !out) return wc_fail("wc_radiomap_jvp: null argument");
    if (!primal->topology || !primal->path_count || primal->topology_capacity <= 0)
        return wc_fail("wc_radiomap_jvp: the forward topology (descriptors and counts) is required");
    if (!primal->field_total || !primal->component_field)
        return wc_fail("wc_radiomap_jvp: the forward fields are required");
    if (!out->d_power_total || !out->d_field_total || !out->d_component_power || !out->d_component_field)
        return wc_fail("wc_radiomap_jvp: tangent outputs are required");
    if (solve->lanes_per_cell < 1 || primal->topology_capacity % solve->lanes_per_cell != 0)
        return wc_fail("wc_radiomap_jvp: topology_capacity must be a multiple of lanes_per_cell");
    if (solve->total_cells <= 0) return 0;
    if (solve->tx_count <= 0) return wc_fail("wc_radiomap_jvp: cells without transmitters");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    const JvpDirection dir{tangent->d_eps_r, tangent->d_sigma_e, tangent->d_frequency, tangent->d_tx_pos};
    const int blocks = (solve->total_cells + 255) / 256;
    hipLaunchKernelGGL(jvp_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, s, *solve, *primal, dir,
                       out->d_power_total, out->d_field_total, out->d_component_power, out->d_component_field,
                       out->d_diffraction_power);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_radiomap_jvp", err);
}
