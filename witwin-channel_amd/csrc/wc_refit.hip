// Scene refit on the device for moved vertices with the same topology (the
// geometry steps of a vertex-gradient loop): triangle tables, the 4-wide BVH's
// quantized boxes, group planes and edge records are recomputed in place; the
// tree's shape, the face groups and the edge list are kept.
//
// Boxes: every wide node is refit after its children (levels deepest first,
// one launch per level). A child's box is its triangle's box (leaf) or the
// union of the child node's decoded boxes, then the node's children are
// re-quantized on the exact grid rule of the host build, so every decoded box
// still contains its subtree. Planes and edges follow the host formulas
// (scene/compiler.py: _group_tables, edge_geometry) in float64.
#include "wc_api_internal.h"
#include "wc_scene.h"
#include "wc_math.h"
#include "../../include/wcmi.h"

#include <cfloat>

namespace wc {
namespace {

__global__ void __launch_bounds__(256) refresh_tris_kernel(const float* __restrict__ vertices,
                                                           const int* __restrict__ faces, int n, Tri* sorted,
                                                           Tri* by_face) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    Tri t = sorted[k];
    const int f = __float_as_int(t.v1.w);
    const int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
    t.v0 = make_float4(vertices[3 * a], vertices[3 * a + 1], vertices[3 * a + 2], t.v0.w);
    t.v1 = make_float4(vertices[3 * b], vertices[3 * b + 1], vertices[3 * b + 2], t.v1.w);
    t.v2 = make_float4(vertices[3 * c], vertices[3 * c + 1], vertices[3 * c + 2], t.v2.w);
    sorted[k] = t;
    by_face[f] = t;
}

struct Box3 {
    float lo[3], hi[3];
};

__device__ inline Box3 decoded_union(const WideNode& w) {
    Box3 b{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    const int count = (int)(w.exps >> 24);
    const float origin[3] = {w.ox, w.oy, w.oz};
    const unsigned qlo[3] = {w.lo_x, w.lo_y, w.lo_z}, qhi[3] = {w.hi_x, w.hi_y, w.hi_z};
    for (int a = 0; a < 3; ++a) {
        const float step = __int_as_float((int)((w.exps >> (8 * a)) & 0xFFu) << 23);
        for (int k = 0; k < count; ++k) {
            b.lo[a] = fminf(b.lo[a], fmaf((float)((qlo[a] >> (8 * k)) & 0xFFu), step, origin[a]));
            b.hi[a] = fmaxf(b.hi[a], fmaf((float)((qhi[a] >> (8 * k)) & 0xFFu), step, origin[a]));
        }
    }
    return b;
}

__device__ inline Box3 tri_box(const Tri& t) {
    Box3 b;
    b.lo[0] = fminf(fminf(t.v0.x, t.v1.x), t.v2.x);
    b.lo[1] = fminf(fminf(t.v0.y, t.v1.y), t.v2.y);
    b.lo[2] = fminf(fminf(t.v0.z, t.v1.z), t.v2.z);
    b.hi[0] = fmaxf(fmaxf(t.v0.x, t.v1.x), t.v2.x);
    b.hi[1] = fmaxf(fmaxf(t.v0.y, t.v1.y), t.v2.y);
    b.hi[2] = fmaxf(fmaxf(t.v0.z, t.v1.z), t.v2.z);
    return b;
}

// The host build's grid rule (wc_bvh_sah.hip: axis_grid): spacing 2^p with the
// extent inside 255 steps and every multiple up to origin + 255 steps exact
// in float.
__device__ inline bool axis_grid(double lo, double hi, int& p_out, double& origin_out) {
    const double extent = hi - lo;
    int p = extent > 0.0 ? (int)ceil(log2(extent / 255.0)) : -126;
    if (p < -126) p = -126;
    for (; p <= 127; ++p) {
        const double step = ldexp(1.0, p);
        const double origin = floor(lo / step) * step;
        const double top = origin + 255.0 * step;
        if (top < hi) continue;
        if (fmax(fabs(origin), fabs(top)) >= ldexp(1.0, p + 24)) continue;
        p_out = p;
        origin_out = origin;
        return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) refit_level_kernel(WideNode* nodes, const int* __restrict__ ids, int count,
                                                          const Tri* __restrict__ sorted, int* failure) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int node = ids[i];
    WideNode w = nodes[node];
    const int kids = (int)(w.exps >> 24);
    const int child[4] = {w.child.x, w.child.y, w.child.z, w.child.w};
    Box3 box[4];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < kids; ++k) {
        box[k] = child[k] < 0 ? tri_box(sorted[~child[k]]) : decoded_union(nodes[child[k]]);
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], (double)box[k].lo[a]);
            hi[a] = fmax(hi[a], (double)box[k].hi[a]);
        }
    }
    double origin[3], step[3];
    unsigned exps = (unsigned)kids << 24;
    for (int a = 0; a < 3; ++a) {
        int p;
        if (!axis_grid(lo[a], hi[a], p, origin[a])) {
            atomicOr(failure, 1);
            return;
        }
        step[a] = ldexp(1.0, p);
        exps |= (unsigned)(p + 127) << (8 * a);
    }
    unsigned qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    for (int k = 0; k < kids; ++k)
        for (int a = 0; a < 3; ++a) {
            const double l = floor(((double)box[k].lo[a] - origin[a]) / step[a]);
            const double h = ceil(((double)box[k].hi[a] - origin[a]) / step[a]);
            if (l < 0.0 || h > 255.0 || l > h) {
                atomicOr(failure, 1);
                return;
            }
            qlo[a] |= (unsigned)l << (8 * k);
            qhi[a] |= (unsigned)h << (8 * k);
        }
    w.ox = (float)origin[0];
    w.oy = (float)origin[1];
    w.oz = (float)origin[2];
    w.exps = exps;
    w.lo_x = qlo[0]; w.lo_y = qlo[1]; w.lo_z = qlo[2];
    w.hi_x = qhi[0]; w.hi_y = qhi[1]; w.hi_z = qhi[2];
    nodes[node] = w;
}

__device__ inline V3<double> vertex_d(const float* v, int i) {
    return {(double)v[3 * i], (double)v[3 * i + 1], (double)v[3 * i + 2]};
}

// Plane of group g from its representative face: unit winding normal with
// its largest-magnitude component made positive, offset n . a.
__global__ void __launch_bounds__(256) plane_kernel(const float* __restrict__ vertices, const int* __restrict__ faces,
                                                    const int* __restrict__ group_face, int groups, double* planes) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    const int f = group_face[g];
    const V3<double> a = vertex_d(vertices, faces[3 * f]), b = vertex_d(vertices, faces[3 * f + 1]),
                     c = vertex_d(vertices, faces[3 * f + 2]);
    V3<double> n = cross(b - a, c - a);
    n = n * (1.0 / fmax(norm(n), 1e-300));
    const double ax = fabs(n.x), ay = fabs(n.y), az = fabs(n.z);
    const double big = ay > ax ? (az > ay ? n.z : n.y) : (az > ax ? n.z : n.x);
    if (big < 0.0) n = n * -1.0;
    planes[4 * g] = n.x;
    planes[4 * g + 1] = n.y;
    planes[4 * g + 2] = n.z;
    planes[4 * g + 3] = dot(n, a);
}

__device__ inline V3<double> unit_or_zero(const V3<double>& v) { return v * (1.0 / fmax(norm(v), 1e-300)); }

// Edge record e from its vertex ids (a, b, c0, c1): origin, direction,
// length, face-0 in-plane direction and exterior normal, wedge index.
__global__ void __launch_bounds__(256) edge_kernel(const float* __restrict__ vertices,
                                                   const int* __restrict__ edge_vertices, int count, Edge* edges) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const int ia = edge_vertices[4 * e], ib = edge_vertices[4 * e + 1], ic0 = edge_vertices[4 * e + 2],
              ic1 = edge_vertices[4 * e + 3];
    const V3<double> A = vertex_d(vertices, ia), B = vertex_d(vertices, ib), C0 = vertex_d(vertices, ic0);
    V3<double> d = B - A;
    const double length = norm(d);
    d = d * (1.0 / fmax(length, 1e-300));
    auto in_plane = [&](const V3<double>& third) {
        V3<double> w = third - A;
        w = w - d * dot(w, d);
        return unit_or_zero(w);
    };
    const V3<double> t0 = in_plane(C0);
    V3<double> n0 = unit_or_zero(cross(B - A, C0 - A));
    n0 = unit_or_zero(n0 - d * dot(n0, d));
    const bool wedge = ic1 >= 0;
    const V3<double> t1 = in_plane(vertex_d(vertices, wedge ? ic1 : ic0));
    double alpha = atan2(-dot(n0, t1), dot(t0, t1));
    if (alpha < 0.0) alpha += kTwoPi;
    Edge r = edges[e];
    r.ax = (float)A.x; r.ay = (float)A.y; r.az = (float)A.z;
    r.dx = (float)d.x; r.dy = (float)d.y; r.dz = (float)d.z;
    r.length = (float)length;
    r.n = (float)(wedge ? 2.0 - alpha / kPi : 2.0);
    r.t0x = (float)t0.x; r.t0y = (float)t0.y; r.t0z = (float)t0.z;
    r.n0x = (float)n0.x; r.n0y = (float)n0.y; r.n0z = (float)n0.z;
    edges[e] = r;
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_scene_refit(const wc_scene* scene, const wc_refit* refit, void* stream) {
    if (!scene || !refit) return wc_fail("wc_scene_refit: null argument");
    if (!refit->vertices || !refit->faces || !refit->failure)
        return wc_fail("wc_scene_refit: vertices, faces and the failure word are required");
    if (scene->face_count < 0 || scene->group_count < 0 || scene->edge_count < 0 || refit->levels < 0)
        return wc_fail("wc_scene_refit: negative count");
    if (scene->face_count > 0 && (!scene->tris || !scene->faces || !scene->nodes || !refit->node_order ||
                                  !refit->level_offset_host || refit->levels <= 0))
        return wc_fail("wc_scene_refit: the triangle tables, nodes and the level order are required");
    if (scene->group_count > 0 && (!refit->group_face || !scene->group_plane))
        return wc_fail("wc_scene_refit: group_face and the group planes are required");
    if (scene->edge_count > 0 && (!refit->edge_vertices || !scene->edges))
        return wc_fail("wc_scene_refit: edge_vertices and the edge records are required");
    hipStream_t st = (hipStream_t)stream;
    const int f = scene->face_count;
    if (f > 0) {
        Tri* sorted = const_cast<Tri*>(static_cast<const Tri*>(scene->tris));
        Tri* by_face = const_cast<Tri*>(static_cast<const Tri*>(scene->faces));
        WideNode* nodes = const_cast<WideNode*>(static_cast<const WideNode*>(scene->nodes));
        hipLaunchKernelGGL(refresh_tris_kernel, dim3((f + 255) / 256), dim3(256), 0, st, refit->vertices,
                           refit->faces, f, sorted, by_face);
        for (int level = 0; level < refit->levels; ++level) {
            const int first = refit->level_offset_host[level], last = refit->level_offset_host[level + 1];
            if (first < 0 || last < first) return wc_fail("wc_scene_refit: level offsets must ascend");
            const int count = last - first;
            if (count == 0) continue;
            hipLaunchKernelGGL(refit_level_kernel, dim3((count + 255) / 256), dim3(256), 0, st, nodes,
                               refit->node_order + first, count, sorted, refit->failure);
        }
    }
    if (scene->group_count > 0)
        hipLaunchKernelGGL(plane_kernel, dim3((scene->group_count + 255) / 256), dim3(256), 0, st, refit->vertices,
                           refit->faces, refit->group_face, scene->group_count,
                           const_cast<double*>(scene->group_plane));
    if (scene->edge_count > 0)
        hipLaunchKernelGGL(edge_kernel, dim3((scene->edge_count + 255) / 256), dim3(256), 0, st, refit->vertices,
                           refit->edge_vertices, scene->edge_count,
                           const_cast<Edge*>(reinterpret_cast<const Edge*>(scene->edges)));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_scene_refit", err);
}
