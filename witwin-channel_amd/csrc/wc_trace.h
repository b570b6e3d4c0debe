// BVH traversal for segment visibility, ordered wall-hit gathering and
// closest-hit discovery rays. Triangle test: watertight ray/triangle
// intersection (Woop, Benthin, Wald, JCGT 2(1), 2013) with the double-precision
// fallback for zero edge functions, so a segment through a shared edge hits
// exactly one of the two triangles and never slips between them.
#pragma once

#include "wc_scene.h"
#include "wc_math.h"

#include <cfloat>

namespace wc {

__device__ __forceinline__ int bits_i(float f) { return __float_as_int(f); }

struct RayPre {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;   // reciprocal direction for slab tests
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float comp(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ RayPre ray_setup(float ox, float oy, float oz, float dx, float dy, float dz) {
    RayPre r;
    r.ox = ox; r.oy = oy; r.oz = oz;
    r.dx = dx; r.dy = dy; r.dz = dz;
    r.ix = 1.0f / dx; r.iy = 1.0f / dy; r.iz = 1.0f / dz;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = 0;
    if (ay > ax) kz = 1;
    if (az > fmaxf(ax, ay)) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (comp(dx, dy, dz, kz) < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    const float dkz = comp(dx, dy, dz, kz);
    r.Sx = comp(dx, dy, dz, kx) / dkz;
    r.Sy = comp(dx, dy, dz, ky) / dkz;
    r.Sz = 1.0f / dkz;
    return r;
}

// Watertight test; returns the hit parameter t in (tmin, tmax) or -1.
__device__ __forceinline__ float tri_hit(const RayPre& r, const Tri& tri, float tmin, float tmax) {
    const float Ax0 = tri.v0.x - r.ox, Ay0 = tri.v0.y - r.oy, Az0 = tri.v0.z - r.oz;
    const float Bx0 = tri.v1.x - r.ox, By0 = tri.v1.y - r.oy, Bz0 = tri.v1.z - r.oz;
    const float Cx0 = tri.v2.x - r.ox, Cy0 = tri.v2.y - r.oy, Cz0 = tri.v2.z - r.oz;
    const float Akz = comp(Ax0, Ay0, Az0, r.kz), Bkz = comp(Bx0, By0, Bz0, r.kz), Ckz = comp(Cx0, Cy0, Cz0, r.kz);
    const float Ax = comp(Ax0, Ay0, Az0, r.kx) - r.Sx * Akz;
    const float Ay = comp(Ax0, Ay0, Az0, r.ky) - r.Sy * Akz;
    const float Bx = comp(Bx0, By0, Bz0, r.kx) - r.Sx * Bkz;
    const float By = comp(Bx0, By0, Bz0, r.ky) - r.Sy * Bkz;
    const float Cx = comp(Cx0, Cy0, Cz0, r.kx) - r.Sx * Ckz;
    const float Cy = comp(Cx0, Cy0, Cz0, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return -1.0f;
    const float det = U + V + W;
    if (det == 0.0f) return -1.0f;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t > tmin && t < tmax)) return -1.0f;
    return t;
}

// Relative widening of the far slab bound. The slab distances and the
// triangle distance are each rounded, so an unwidened box test can reject a
// box whose triangle the watertight test hits (a ray grazing the rim of a
// flat box at a shared edge or corner). Widening the far bound by a few
// ulps (Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013, uses 1 + 2 gamma_3;
// 32 ulps also covers the rounding of the triangle distance short of
// near-parallel hits) lets every walk order reach the same triangles, so
// any-hit and gathered answers do not depend on the tree. The triangle test's
// distance carries an absolute error of a few ulps of the vertex coordinates
// (a large triangle far from the segment's origin): closest hits on two
// surfaces that meet within that error along the segment may resolve either
// way, as they would under a different rounding.
constexpr float kSlabWiden = 1.0f + 32.0f * FLT_EPSILON;

__device__ __forceinline__ bool slab(const RayPre& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float tmin, float tmax, float& tnear) {
    const float t0x = (lx - r.ox) * r.ix, t1x = (hx - r.ox) * r.ix;
    const float t0y = (ly - r.oy) * r.iy, t1y = (hy - r.oy) * r.iy;
    const float t0z = (lz - r.oz) * r.iz, t1z = (hz - r.oz) * r.iz;
    const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    // only the box's own bound is widened: the segment's end bound stays exact
    const float tf = fminf(fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z)) * kSlabWiden, tmax);
    tnear = tn;
    return tn <= tf;
}

__device__ __forceinline__ float grid_step(unsigned exps, int axis) {
    return __int_as_float((int)((exps >> (8 * axis)) & 0xFFu) << 23);
}

// bound of child k on one axis: origin + q * step, exact by construction
__device__ __forceinline__ float grid_bound(unsigned packed, int k, float step, float origin) {
    return fmaf((float)((packed >> (8 * k)) & 0xFFu), step, origin);
}

// The tree's top levels (wide nodes 0 .. kTopNodes - 1, numbered breadth-first
// by the host build) staged in LDS. Every walk opens at the root and a
// wavefront's segments share their first levels, so these fetches repeat in
// every lane of every walk; from LDS they skip the vector memory path. Only a
// kernel that calls stage_top_nodes before its first walk may walk with
// kTop = true: the array holds nothing until then.
__shared__ uint4 wc_top_words[kTopNodes * 4];

// Cooperative copy of the top nodes into LDS; every thread of the workgroup
// must reach it (it ends in a barrier), before any early return.
__device__ __forceinline__ void stage_top_nodes(const SceneView& s) {
    const int words = (s.node_count < kTopNodes ? s.node_count : kTopNodes) * 4;
    const uint4* src = reinterpret_cast<const uint4*>(s.nodes);
    for (int i = (int)threadIdx.x; i < words; i += (int)blockDim.x) wc_top_words[i] = src[i];
    __syncthreads();
}

template <bool kTop>
__device__ __forceinline__ WideNode fetch_node(const SceneView& s, int node) {
    if (kTop && node < kTopNodes) {
        WideNode nd;
        uint4* w = reinterpret_cast<uint4*>(&nd);
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = wc_top_words[4 * node + i];
        return nd;
    }
    return s.nodes[node];
}

// Walks the 4-wide tree over the segment (tmin, limit), handing every leaf
// whose box it crosses to `leaf` (which returns true to end the walk). Of the
// internal children a node's segment crosses, the nearest is entered next and
// the others wait on the stack, so a blocker close to the segment's start ends
// an any-hit walk early; a node's leaves are tested after its box tests, one
// triangle per lane per round. `limit` is read at every box test: a
// closest-hit leaf may shrink it. The stack never overflows: the host build
// checks the tree's worst-case stack need against kWideStackSize.
template <bool kTop = false, class Leaf>
__device__ __forceinline__ void walk_wide(const SceneView& s, const RayPre& r, float tmin, const float& limit,
                                          Leaf&& leaf) {
    int stack[kWideStackSize];
    int sp = 0;
    int node = 0;
    while (true) {
        const WideNode nd = fetch_node<kTop>(s, node);
        const float sx = grid_step(nd.exps, 0), sy = grid_step(nd.exps, 1), sz = grid_step(nd.exps, 2);
        const int count = (int)(nd.exps >> 24);
        int next = -1;
        float tnext = 0.0f;
        unsigned leaves = 0;  // slots of crossed leaves, tested after the box tests
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float tn;
            if (k < count &&
                slab(r, grid_bound(nd.lo_x, k, sx, nd.ox), grid_bound(nd.lo_y, k, sy, nd.oy),
                     grid_bound(nd.lo_z, k, sz, nd.oz), grid_bound(nd.hi_x, k, sx, nd.ox),
                     grid_bound(nd.hi_y, k, sy, nd.oy), grid_bound(nd.hi_z, k, sz, nd.oz), tmin, limit, tn)) {
                const int c = k == 0 ? nd.child.x : (k == 1 ? nd.child.y : (k == 2 ? nd.child.z : nd.child.w));
                if (c < 0) {
                    leaves |= 1u << k;
                } else if (next < 0) {
                    next = c;
                    tnext = tn;
                } else {
                    int wait = c;
                    if (tn < tnext) {
                        wait = next;
                        next = c;
                        tnext = tn;
                    }
                    if (sp < kWideStackSize) stack[sp++] = wait;
                }
            }
        }
        // one triangle per lane per round: lanes whose leaves sit in different
        // slots share the rounds instead of taking turns slot by slot
        while (leaves) {
            const int k = __ffs(leaves) - 1;
            leaves &= leaves - 1;
            const int c = k == 0 ? nd.child.x : (k == 1 ? nd.child.y : (k == 2 ? nd.child.z : nd.child.w));
            if (leaf(s.tris[~c])) return;
        }
        if (next < 0) {
            if (sp == 0) return;
            next = stack[--sp];
        }
        node = next;
    }
}

// Any hit on the segment o + t*d, t in (tmin, tmax), skipping triangles of
// groups ig0/ig1 (the surfaces the segment starts or ends on).
template <bool kTop = false>
__device__ inline bool occluded(const SceneView& s, const RayPre& r, float tmin, float tmax, int ig0, int ig1) {
    if (s.tri_count == 0) return false;
    bool blocked = false;
    walk_wide<kTop>(s, r, tmin, tmax, [&](const Tri& tri) {
        const int g = bits_i(tri.v0.w);
        blocked = g != ig0 && g != ig1 && tri_hit(r, tri, tmin, tmax) >= 0.0f;
        return blocked;
    });
    return blocked;
}

struct Hit {
    float t;
    int face;
    int group;
};

// Closest hit in (tmin, tmax), skipping group `ignore`; t < 0 when nothing is
// hit. Equal hit distances (a ray through an edge or corner shared by two
// faces) resolve to the lower face id, so the answer does not depend on the
// order in which the tree is walked (the widened box test, with the limit
// itself widened once a hit exists, reaches every tied triangle), up to the
// rounding of the triangle distance noted at kSlabWiden.
__device__ inline void closest_candidate(const RayPre& r, const Tri& tri, float tmin, int ignore, float& tbest,
                                         Hit& best) {
    const int g = bits_i(tri.v0.w);
    if (g == ignore) return;
    const float lim = best.face >= 0 ? nextafterf(tbest, INFINITY) : tbest;
    const float t = tri_hit(r, tri, tmin, lim);
    if (t < 0.0f) return;
    const int f = bits_i(tri.v1.w);
    if (t < tbest || f < best.face) {
        tbest = t;
        best = {t, f, g};
    }
}

template <bool kTop = false>
__device__ inline Hit closest_hit(const SceneView& s, const RayPre& r, float tmin, float tmax, int ignore) {
    Hit best{-1.0f, -1, -1};
    if (s.tri_count == 0) return best;
    float tbest = tmax;
    float limit = tmax;
    walk_wide<kTop>(s, r, tmin, limit, [&](const Tri& tri) {
        closest_candidate(r, tri, tmin, ignore, tbest, best);
        // once a hit exists, boxes up to a few ulps past it still hold ties
        limit = best.face >= 0 ? tbest * kSlabWiden : tbest;
        return false;
    });
    return best;
}

// Order of gathered hits: distance, then group id for equal distances (a
// segment through an edge shared by two walls), independent of walk order.
__device__ __forceinline__ bool hit_before(const Hit& a, const Hit& b) {
    return a.t < b.t || (a.t == b.t && a.group < b.group);
}

// All hits on the segment in (tmin, tmax), ordered by t, at most one per
// coplanar group (a segment crossing a wall through the shared diagonal of
// its two triangles counts once, with the group's lowest face id). Returns
// the number of distinct hits, or cap + 1 when there are more than `cap`.
template <bool kTop = false>
__device__ inline int gather_hits(const SceneView& s, const RayPre& r, float tmin, float tmax, int cap,
                                  Hit* hits) {
    int count = 0;
    if (s.tri_count == 0) return 0;
    walk_wide<kTop>(s, r, tmin, tmax, [&](const Tri& tri) {
        const float t = tri_hit(r, tri, tmin, tmax);
        if (t < 0.0f) return false;
        const int g = bits_i(tri.v0.w);
        const int f = bits_i(tri.v1.w);
        bool dup = false;
        for (int i = 0; i < count && i < cap; ++i) {
            if (hits[i].group != g) continue;
            dup = true;
            if (f < hits[i].face) {
                // the group's lowest face and its distance; restore the order
                hits[i].face = f;
                hits[i].t = t;
                int j = i;
                while (j > 0 && hit_before(hits[j], hits[j - 1])) {
                    const Hit tmp = hits[j]; hits[j] = hits[j - 1]; hits[j - 1] = tmp; --j;
                }
                while (j + 1 < count && j + 1 < cap && hit_before(hits[j + 1], hits[j])) {
                    const Hit tmp = hits[j]; hits[j] = hits[j + 1]; hits[j + 1] = tmp; ++j;
                }
            }
        }
        if (dup) return false;
        if (count < cap) {
            // insertion by (t, group)
            const Hit h{t, f, g};
            int j = count;
            while (j > 0 && hit_before(h, hits[j - 1])) { hits[j] = hits[j - 1]; --j; }
            hits[j] = h;
        }
        ++count;
        return count > cap;
    });
    return count > cap ? cap + 1 : count;
}

}  // namespace wc
