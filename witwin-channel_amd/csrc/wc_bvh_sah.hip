// Surface-area-heuristic BVH, built on the host for static scenes.
//
// Top-down binned SAH (16 centroid bins per axis, one triangle per leaf) with
// the same output layout as the device LBVH: node 0 is the root, each node
// holds both child boxes inline, a negative child ~k names triangle k of the
// leaf-ordered table. Against the LBVH it cuts the nodes an occlusion walk
// visits by about a third on building and terrain meshes (long wall and
// roof triangles spoil Morton order), which every segment test of the
// radiomap pays. Equal split costs keep the first candidate of a fixed scan
// order and partitions are stable, so the tree is a pure function of the
// input arrays.
#include "wc_api_internal.h"
#include "wc_scene.h"

#include <cfloat>
#include <cstring>
#include <vector>

namespace wc {
namespace {

constexpr int kSahBins = 16;
// Subtrees that would reach this depth fall back to median splits, which
// add at most ceil(log2(count)) levels: the tree stays shallower than the
// traversal stack.
constexpr int kSahDepthLimit = kStackSize - 4;

struct HostBox {
    float lo[3], hi[3];
};

inline HostBox empty_box() {
    HostBox b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = FLT_MAX;
        b.hi[k] = -FLT_MAX;
    }
    return b;
}

inline void grow(HostBox& b, const HostBox& o) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = b.lo[k] < o.lo[k] ? b.lo[k] : o.lo[k];
        b.hi[k] = b.hi[k] > o.hi[k] ? b.hi[k] : o.hi[k];
    }
}

inline double half_area(const HostBox& b) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = b.hi[k] > b.lo[k] ? (double)b.hi[k] - (double)b.lo[k] : 0.0;
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

inline float int_bits(int v) {
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

inline int ceil_log2(int n) {
    int d = 0;
    while ((1 << d) < n) ++d;
    return d;
}

struct Task {
    int begin, end, node, depth;
};

struct Builder {
    const std::vector<HostBox>& box;
    const std::vector<float>& cen;  // [n, 3]
    std::vector<int>& idx;
    std::vector<int> scratch;

    int bin_of(int prim, int axis, float lo, float scale) const {
        int b = (int)((cen[3 * prim + axis] - lo) * scale);
        return b < 0 ? 0 : (b >= kSahBins ? kSahBins - 1 : b);
    }

    HostBox bounds(int begin, int end) const {
        HostBox b = empty_box();
        for (int i = begin; i < end; ++i) grow(b, box[idx[i]]);
        return b;
    }

    // Split [begin, end) in place; returns the first index of the right part.
    int split(int begin, int end, bool median) {
        const int count = end - begin;
        HostBox cb = empty_box();
        for (int i = begin; i < end; ++i)
            for (int k = 0; k < 3; ++k) {
                const float c = cen[3 * idx[i] + k];
                cb.lo[k] = c < cb.lo[k] ? c : cb.lo[k];
                cb.hi[k] = c > cb.hi[k] ? c : cb.hi[k];
            }
        if (!median) {
            double best = 1e300;
            int best_axis = -1, best_bin = 0;
            for (int axis = 0; axis < 3; ++axis) {
                const float ext = cb.hi[axis] - cb.lo[axis];
                if (!(ext > 0.0f)) continue;
                const float scale = (float)kSahBins / ext;
                HostBox bins[kSahBins];
                int counts[kSahBins] = {};
                for (int b = 0; b < kSahBins; ++b) bins[b] = empty_box();
                for (int i = begin; i < end; ++i) {
                    const int b = bin_of(idx[i], axis, cb.lo[axis], scale);
                    grow(bins[b], box[idx[i]]);
                    ++counts[b];
                }
                double left_cost[kSahBins];
                HostBox acc = empty_box();
                int n = 0;
                for (int b = 0; b < kSahBins - 1; ++b) {
                    grow(acc, bins[b]);
                    n += counts[b];
                    left_cost[b] = n ? half_area(acc) * n : 0.0;
                }
                acc = empty_box();
                n = 0;
                for (int b = kSahBins - 1; b >= 1; --b) {
                    grow(acc, bins[b]);
                    n += counts[b];
                    if (n == 0 || n == count) continue;
                    const double cost = left_cost[b - 1] + half_area(acc) * n;
                    // strict improvement only: an equal cost keeps the earlier
                    // (axis, bin) of the scan
                    if (cost < best) {
                        best = cost;
                        best_axis = axis;
                        best_bin = b;
                    }
                }
            }
            if (best_axis >= 0) {
                const float scale = (float)kSahBins / (cb.hi[best_axis] - cb.lo[best_axis]);
                // stable partition through the scratch buffer
                int left = begin;
                int right = 0;
                for (int i = begin; i < end; ++i) {
                    const int p = idx[i];
                    if (bin_of(p, best_axis, cb.lo[best_axis], scale) < best_bin) idx[left++] = p;
                    else scratch[right++] = p;
                }
                std::memcpy(&idx[left], scratch.data(), sizeof(int) * right);
                return left;
            }
        }
        // median split along the widest centroid extent (stable order, ties by
        // position in the current range)
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
        const int mid = begin + count / 2;
        std::vector<int> order(idx.begin() + begin, idx.begin() + end);
        std::vector<int> rank(count);
        for (int i = 0; i < count; ++i) rank[i] = i;
        std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) {
            return cen[3 * order[a] + axis] < cen[3 * order[b] + axis];
        });
        for (int i = 0; i < count; ++i) idx[begin + i] = order[rank[i]];
        return mid;
    }
};

void write_child(Node& nd, int which, const HostBox& b, int child) {
    if (which == 0) {
        nd.a = make_float4(b.lo[0], b.lo[1], b.lo[2], b.hi[0]);
        nd.b.x = b.hi[1];
        nd.b.y = b.hi[2];
        nd.d.x = child;
    } else {
        nd.b.z = b.lo[0];
        nd.b.w = b.lo[1];
        nd.c = make_float4(b.lo[2], b.hi[0], b.hi[1], b.hi[2]);
        nd.d.y = child;
    }
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_bvh_build_sah(const float* vertices, const int* faces, const int* face_group,
                                const int* face_material, int face_count, void* nodes_out, void* sorted_tris_out,
                                void* face_tris_out) {
    if (face_count < 0) return wc_fail("wc_bvh_build_sah: face_count must be non-negative");
    if (face_count == 0) return 0;
    if (!vertices || !faces || !face_group || !face_material || !nodes_out || !sorted_tris_out || !face_tris_out)
        return wc_fail("wc_bvh_build_sah: null pointer argument");
    const int n = face_count;
    Tri* sorted = static_cast<Tri*>(sorted_tris_out);
    Tri* by_face = static_cast<Tri*>(face_tris_out);
    Node* nodes = static_cast<Node*>(nodes_out);
    std::vector<HostBox> box(n);
    std::vector<float> cen(3 * (size_t)n);
    for (int f = 0; f < n; ++f) {
        const int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
        if (a < 0 || b < 0 || c < 0) return wc_fail("wc_bvh_build_sah: negative vertex index in face %d", f);
        Tri t;
        t.v0 = make_float4(vertices[3 * a], vertices[3 * a + 1], vertices[3 * a + 2], int_bits(face_group[f]));
        t.v1 = make_float4(vertices[3 * b], vertices[3 * b + 1], vertices[3 * b + 2], int_bits(f));
        t.v2 = make_float4(vertices[3 * c], vertices[3 * c + 1], vertices[3 * c + 2], int_bits(face_material[f]));
        by_face[f] = t;
        const float xs[3] = {t.v0.x, t.v1.x, t.v2.x}, ys[3] = {t.v0.y, t.v1.y, t.v2.y},
                    zs[3] = {t.v0.z, t.v1.z, t.v2.z};
        const float* comps[3] = {xs, ys, zs};
        for (int k = 0; k < 3; ++k) {
            const float* v = comps[k];
            box[f].lo[k] = fminf(fminf(v[0], v[1]), v[2]);
            box[f].hi[k] = fmaxf(fmaxf(v[0], v[1]), v[2]);
            cen[3 * (size_t)f + k] = (v[0] + v[1] + v[2]) * (1.0f / 3.0f);
        }
    }
    if (n == 1) {
        sorted[0] = by_face[0];
        Node nd;
        const HostBox& b = box[0];
        nd.a = make_float4(b.lo[0], b.lo[1], b.lo[2], b.hi[0]);
        // second child: inverted (empty) box never hit
        nd.b = make_float4(b.hi[1], b.hi[2], FLT_MAX, FLT_MAX);
        nd.c = make_float4(FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX);
        nd.d = make_int4(~0, ~0, 0, 0);
        nodes[0] = nd;
        return 0;
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    Builder bld{box, cen, idx, std::vector<int>(n)};
    int next_node = 1;
    std::vector<Task> stack;
    stack.push_back({0, n, 0, 0});
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        const int count = t.end - t.begin;
        const bool median = t.depth + ceil_log2(count) >= kSahDepthLimit;
        const int mid = bld.split(t.begin, t.end, median);
        Node nd;
        nd.d = make_int4(0, 0, 0, 0);
        const int parts[2][2] = {{t.begin, mid}, {mid, t.end}};
        for (int side = 0; side < 2; ++side) {
            const int b = parts[side][0], e = parts[side][1];
            const HostBox cb = bld.bounds(b, e);
            int child;
            if (e - b == 1) {
                child = ~b;  // leaf: triangle at leaf position b
            } else {
                child = next_node++;
                stack.push_back({b, e, child, t.depth + 1});
            }
            write_child(nd, side, cb, child);
        }
        nodes[t.node] = nd;
    }
    if (next_node != n - 1) return wc_fail("wc_bvh_build_sah: internal node count %d != %d", next_node, n - 1);
    for (int i = 0; i < n; ++i) sorted[i] = by_face[idx[i]];
    return 0;
}
