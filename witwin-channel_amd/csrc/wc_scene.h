// Device-resident scene layout shared by the BVH builder, the tracers and the
// path kernels. All views are plain pointers into HBM owned by the caller
// (torch tensors on the Python side); the library allocates only the BVH.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace wc {

constexpr int kMaxDepth = 5;          // public reflection/transmission depth cap
constexpr int kMaxHits = 8;           // ordered wall hits gathered on a straight segment
constexpr int kStackSize = 64;        // depth bound of the binary trees the builders produce
constexpr int kWideStackSize = 96;    // traversal stack of the 4-wide tree (checked at build time)
// The 4-wide tree's first kTopNodes nodes in breadth-first order from the
// root (its top five levels when they are full) take ids 0 .. kTopNodes - 1;
// the walk kernels stage them in LDS (21,824 B per workgroup). Scaled-city
// chain validity pass per batch: 50.7 ms with no staging, 50.8 / 50.3 /
// 50.0 ms staging 21 / 85 / 341 nodes.
#ifndef WC_TOP_NODES
#define WC_TOP_NODES 341
#endif
constexpr int kTopNodes = WC_TOP_NODES;

// Triangle in BVH (Morton) order for intersection: three float4 rows,
// v0.w = group id, v1.w = original face id, v2.w = material id (bit casts).
struct Tri {
    float4 v0, v1, v2;
};

// Binary LBVH node with both child boxes inline (one 64 B load tests two children).
// child >= 0: internal node index; child < 0: leaf holding sorted triangle ~child.
struct Node {
    float4 a;  // c0.lo.x c0.lo.y c0.lo.z c0.hi.x
    float4 b;  // c0.hi.y c0.hi.z c1.lo.x c1.lo.y
    float4 c;  // c1.lo.z c1.hi.x c1.hi.y c1.hi.z
    int4 d;    // child0 child1 - -
};

// Wedge (two faces) or half-plane (one face) diffraction edge. Angles are
// measured in the plane normal to `dir`, from `t0` (into face 0) towards the
// exterior side given by `n0`; the exterior spans [0, n*pi].
struct Edge {
    float ax, ay, az;      // start point
    float dx, dy, dz;      // unit direction
    float length;          // edge length (m)
    float n;               // exterior wedge index n (2 for a half-plane)
    float t0x, t0y, t0z;   // unit in-plane direction into face 0, perpendicular to dir
    float n0x, n0y, n0z;   // face 0 normal, pointing into the exterior
    int face0, face1;      // original face ids (face1 = -1 for a half-plane)
    int group0, group1;    // coplanar groups of the faces (-1 when absent)
    int material0, material1;
};

// 4-wide node the tracers walk: 64 B, one cache line. Child boxes are 8-bit
// offsets on a per-axis grid of spacing 2^(e - 127) from `origin`, a multiple
// of that spacing; the host picks e so that every decoded bound
// origin + q * 2^(e - 127) (q = 0..255) is a float exactly, hence a decoded
// box always contains the child's true box. child >= 0: wide node index;
// child < 0: leaf holding sorted triangle ~child. Slots k >= the child count
// are unused. Built on the host from a binary tree (wc_bvh_widen).
struct WideNode {
    float ox, oy, oz;
    unsigned exps;              // bytes 0..2: biased exponents e of x, y, z; byte 3: child count (1..4)
    unsigned lo_x, lo_y, lo_z;  // byte k: child k's lower bound in grid steps
    unsigned hi_x, hi_y, hi_z;  // byte k: child k's upper bound in grid steps
    unsigned pad0, pad1;
    int4 child;
};
static_assert(sizeof(WideNode) == 64, "a wide node fills one 64-byte cache line");

struct SceneView {
    // 4-wide BVH over triangles in sorted order
    const WideNode* nodes;
    const Tri* tris;
    int tri_count;
    int node_count;
    // original face order
    const Tri* faces;        // same packing as Tri, indexed by face id
    const int* face_group;   // face -> coplanar group
    const int* face_material;
    int face_count;
    // coplanar reflection groups
    const double* group_plane;  // [G,4] unit normal xyz + offset (n.x = d), double precision
    const int* group_face_offset;  // [G+1]
    const int* group_faces;        // faces of each group, ascending
    const int* group_material;     // material of the group's representative face
    int group_count;
    // materials (scalar first-layer view + CSR layer stacks)
    const float* mat_eps;
    const float* mat_sigma;
    const float* mat_mu;
    const float* mat_gain;
    const float* mat_thickness;
    const int* mat_layer_offset;
    const int* mat_layer_count;
    const float* layer_thickness;
    const float* layer_eps;
    const float* layer_sigma;
    const float* layer_mu;
    int material_count;
    // diffraction edges
    const Edge* edges;
    int edge_count;
};

}  // namespace wc
