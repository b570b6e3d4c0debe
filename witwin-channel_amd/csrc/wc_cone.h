// Receiver cones of reflection chains: a conservative bound of the receivers
// for which a chain's back-projection can land inside every group it bounces
// on (beam tracing over the convex hulls of the groups).
//
// For a chain of groups g_1..g_D with image points I_1..I_D, the hit h_k on
// group k lies in the polygon V_k: V_1 is the hull of g_1, and V_k is the
// hull of g_k clipped to the cone from I_{k-1} through V_{k-1} (the points
// x = I_{k-1} + t (p - I_{k-1}), p in V_{k-1}, t >= 1: beyond plane k-1).
// A receiver r can only be valid when r lies in the cone from I_D through
// V_D, beyond plane D. Every polygon is grown by a margin before it is used
// and every plane is shifted outward by it, so the bound also covers the
// exact test's float rounding (its inclusive point-in-triangle test on the
// float-rounded hit).
//
// The cone is stored relative to a local origin o (the centroid of V_D, on
// plane D) as planes n . (x - o) >= w with unit n: w is small (each side
// plane contains an edge of V_D near o), so float storage keeps the
// distances accurate to far below the margin.
#pragma once

#include "wc_math.h"
#include "wc_scene.h"

namespace wc {

constexpr int kConePlanes = 11;     // side planes (edges of V_D) + plane D
constexpr int kConeFloats = 4 + 4 * kConePlanes;  // origin (xyz, plane count) + planes
constexpr int kConePoly = 16;       // polygon vertices kept while clipping

struct ConePoly {
    V3<double> v[kConePoly];
    int n;
};

// grow a convex polygon in the plane with unit normal nrm by distance m
// (each edge moved outward by m; a vertex moves to the meeting point of its
// two moved edges, capped for very sharp corners)
__host__ __device__ inline void grow_polygon(ConePoly& p, const V3<double>& nrm, double m) {
    if (p.n < 3) return;
    V3<double> out[kConePoly];
    // orientation: outward in-plane normal of edge a->b is (b - a) x nrm when
    // the polygon runs counter-clockwise about nrm
    V3<double> area{0.0, 0.0, 0.0};
    for (int i = 0; i < p.n; ++i) area = area + cross(p.v[i], p.v[(i + 1) % p.n]);
    const double sgn = dot(area, nrm) >= 0.0 ? 1.0 : -1.0;
    for (int i = 0; i < p.n; ++i) {
        const V3<double> a = p.v[(i + p.n - 1) % p.n], b = p.v[i], c = p.v[(i + 1) % p.n];
        V3<double> n1 = cross(b - a, nrm) * sgn, n2 = cross(c - b, nrm) * sgn;
        const double l1 = norm(n1), l2 = norm(n2);
        if (l1 <= 0.0 || l2 <= 0.0) {
            out[i] = b;
            continue;
        }
        n1 = n1 * (1.0 / l1);
        n2 = n2 * (1.0 / l2);
        const double den = fmax(1.0 + dot(n1, n2), 1.0e-3);
        out[i] = b + (n1 + n2) * (m / den);
    }
    for (int i = 0; i < p.n; ++i) p.v[i] = out[i];
}

// keep the part of p with dot(n, x) - d >= -m (Sutherland-Hodgman); false
// when more than kConePoly vertices would be needed (caller gives up)
__host__ __device__ inline bool clip_polygon(ConePoly& p, const V3<double>& n, double d, double m) {
    if (p.n == 0) return true;
    ConePoly r;
    r.n = 0;
    for (int i = 0; i < p.n; ++i) {
        const V3<double> a = p.v[i], b = p.v[(i + 1) % p.n];
        const double fa = dot(n, a) - d + m, fb = dot(n, b) - d + m;
        if (fa >= 0.0) {
            if (r.n >= kConePoly) return false;
            r.v[r.n++] = a;
        }
        if ((fa >= 0.0) != (fb >= 0.0)) {
            if (r.n >= kConePoly) return false;
            const double t = fa / (fa - fb);
            r.v[r.n++] = a + (b - a) * t;
        }
    }
    p = r;
    return true;
}

// side planes of the cone from apex I through convex polygon V (beyond V):
// for edge (a, b) the plane through I, a, b with the polygon on its positive
// side; calls emit(n, d) with unit n, plane dot(n, x) = d
template <class Emit>
__host__ __device__ inline bool cone_sides(const ConePoly& V, const V3<double>& I, Emit&& emit) {
    V3<double> c{0.0, 0.0, 0.0};
    for (int i = 0; i < V.n; ++i) c = c + V.v[i];
    c = c * (1.0 / V.n);
    for (int i = 0; i < V.n; ++i) {
        const V3<double> a = V.v[i], b = V.v[(i + 1) % V.n];
        V3<double> n = cross(a - I, b - I);
        const double l = norm(n);
        if (!(l > 0.0)) continue;  // degenerate edge: no plane (looser cone)
        n = n * (1.0 / l);
        double d = dot(n, I);
        if (dot(n, c) - d < 0.0) {
            n = -n;
            d = -d;
        }
        emit(n, d);
    }
    return true;
}

// Cone of one chain (groups[0..D-1], images[D][3]: the transmitter mirrored
// through groups 0..k); writes kConeFloats floats, out[3] = number of planes,
// 0 = no bound (the chain is never culled by it); a chain whose region is
// empty gets one plane no point satisfies.
// hull_offset [G + 1] / hull_points [H][3]: convex hull of each group (in
// its plane). plane(g) = (n, off) with unit n (group_plane).
__host__ __device__ inline void chain_cone(const double* __restrict__ group_plane, const int* __restrict__ hull_offset,
                                           const double* __restrict__ hull_points, const int* groups,
                                           const double* images, int D, float* out) {
    for (int i = 0; i < kConeFloats; ++i) out[i] = 0.0f;
    if (D < 1 || D > kMaxDepth) return;
    ConePoly V;
    V.n = 0;
    for (int k = 0; k < D; ++k) {
        const int g = groups[k];
        const V3<double> nrm{group_plane[4 * g], group_plane[4 * g + 1], group_plane[4 * g + 2]};
        const int h0 = hull_offset[g], h1 = hull_offset[g + 1];
        if (h1 - h0 < 3 || h1 - h0 > kConePoly / 2) return;
        double scale = 0.0;
        ConePoly P;
        P.n = h1 - h0;
        for (int i = 0; i < P.n; ++i) {
            P.v[i] = {hull_points[3 * (h0 + i)], hull_points[3 * (h0 + i) + 1], hull_points[3 * (h0 + i) + 2]};
            scale = fmax(scale, fabs(P.v[i].x) + fabs(P.v[i].y) + fabs(P.v[i].z));
        }
        const double m = 1.0e-3 + 1.0e-5 * scale;
        grow_polygon(P, nrm, m);
        if (k > 0) {
            // the previous hit polygon seen from the previous image, beyond
            // the previous plane
            const V3<double> I{images[3 * (k - 1)], images[3 * (k - 1) + 1], images[3 * (k - 1) + 2]};
            const int gp = groups[k - 1];
            V3<double> np{group_plane[4 * gp], group_plane[4 * gp + 1], group_plane[4 * gp + 2]};
            double dp = group_plane[4 * gp + 3];
            if (dot(np, I) - dp > 0.0) {  // beyond = the side away from the image
                np = -np;
                dp = -dp;
            }
            if (fabs(dot(np, I) - dp) < 1.0e-6) return;  // image on its plane: no bound
            bool ok = clip_polygon(P, np, dp, m);
            cone_sides(V, I, [&](const V3<double>& n, double d) { ok = ok && clip_polygon(P, n, d, m); });
            if (!ok) return;
        }
        if (P.n < 3) {
            // empty even with every margin: no receiver validates the chain;
            // one plane that no point satisfies (0 >= 1)
            out[3] = 1.0f;
            out[7] = 1.0f;
            return;
        }
        V = P;
    }
    // receiver cone: apex I_D through V_D, beyond plane D
    const V3<double> I{images[3 * (D - 1)], images[3 * (D - 1) + 1], images[3 * (D - 1) + 2]};
    const int g = groups[D - 1];
    V3<double> nd{group_plane[4 * g], group_plane[4 * g + 1], group_plane[4 * g + 2]};
    double dd = group_plane[4 * g + 3];
    if (dot(nd, I) - dd > 0.0) {
        nd = -nd;
        dd = -dd;
    }
    if (fabs(dot(nd, I) - dd) < 1.0e-6) return;
    V3<double> o{0.0, 0.0, 0.0};
    double scale = 0.0;
    for (int i = 0; i < V.n; ++i) {
        o = o + V.v[i];
        scale = fmax(scale, fabs(V.v[i].x) + fabs(V.v[i].y) + fabs(V.v[i].z));
    }
    o = o * (1.0 / V.n);
    const double m = 1.0e-3 + 1.0e-5 * scale;
    int planes = 0;
    auto put = [&](const V3<double>& n, double d) {
        if (planes >= kConePlanes) return;  // dropping a plane only loosens the bound
        float* p = out + 4 + 4 * planes;
        p[0] = (float)n.x;
        p[1] = (float)n.y;
        p[2] = (float)n.z;
        p[3] = (float)(d - dot(n, o) - m);  // n . (x - o) >= w, shifted out by m
        ++planes;
    };
    put(nd, dd);
    cone_sides(V, I, put);
    out[0] = (float)o.x;
    out[1] = (float)o.y;
    out[2] = (float)o.z;
    out[3] = (float)planes;
}

// false when no point of the box [lo, hi] lies in the cone (every corner
// outside one plane by more than the float slack of the test)
__host__ __device__ __forceinline__ bool cone_reaches_box(const float* __restrict__ cone, const V3<float>& lo,
                                                          const V3<float>& hi) {
    const int planes = (int)cone[3];
    if (planes <= 0) return true;
    const V3<float> o{cone[0], cone[1], cone[2]};
    const V3<float> c{0.5f * (lo.x + hi.x) - o.x, 0.5f * (lo.y + hi.y) - o.y, 0.5f * (lo.z + hi.z) - o.z};
    const V3<float> h{0.5f * (hi.x - lo.x), 0.5f * (hi.y - lo.y), 0.5f * (hi.z - lo.z)};
    // float slack of n . (x - o) for |x - o| up to the box's reach
    const float slack = 1.0e-6f * (fabsf(c.x) + fabsf(c.y) + fabsf(c.z) + h.x + h.y + h.z) + 1.0e-5f;
    for (int i = 0; i < planes && i < kConePlanes; ++i) {
        const float* p = cone + 4 + 4 * i;
        const float best = p[0] * c.x + p[1] * c.y + p[2] * c.z + fabsf(p[0]) * h.x + fabsf(p[1]) * h.y +
                           fabsf(p[2]) * h.z;
        if (best < p[3] - slack) return false;
    }
    return true;
}

}  // namespace wc
