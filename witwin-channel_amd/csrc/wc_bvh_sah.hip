// Host-side BVH work: the surface-area-heuristic build for static scenes and
// the collapse of any binary tree into the 4-wide tree the tracers walk.
//
// SAH build: top-down binned SAH (16 centroid bins per axis, one triangle per
// leaf) into the binary layout of the device LBVH (node 0 the root, both
// child boxes inline, a negative child ~k names triangle k of the
// leaf-ordered table). Against the LBVH it cuts the nodes an occlusion walk
// visits by about a third on building and terrain meshes (long wall and roof
// triangles spoil Morton order). Equal split costs keep the first candidate of
// a fixed scan order and partitions are stable, so the tree is a pure
// function of the input arrays.
//
// Collapse: every wide node takes a binary node's two children and keeps
// opening the internal child with the largest surface area until it holds
// four, which halves the dependent node fetches of a walk; its child boxes are
// stored as 8-bit steps of an exact per-axis grid (64 B per node). The walk keeps the
// unvisited siblings of its path on a stack; the collapse computes the
// tree's exact worst case and, should it exceed the tracers' stack, redoes
// the collapse opening every internal child once (depth halves, so a binary
// tree of depth <= 60 needs at most 90 entries).
#include "wc_api_internal.h"
#include "wc_scene.h"

#include <cfloat>
#include <cmath>
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

struct ChildRef {
    int child;
    HostBox box;
};

inline HostBox node_child_box(const Node& nd, int which) {
    HostBox b;
    if (which == 0) {
        b.lo[0] = nd.a.x; b.lo[1] = nd.a.y; b.lo[2] = nd.a.z;
        b.hi[0] = nd.a.w; b.hi[1] = nd.b.x; b.hi[2] = nd.b.y;
    } else {
        b.lo[0] = nd.b.z; b.lo[1] = nd.b.w; b.lo[2] = nd.c.x;
        b.hi[0] = nd.c.y; b.hi[1] = nd.c.z; b.hi[2] = nd.c.w;
    }
    return b;
}

// Grid of one axis: the spacing 2^p and an origin that is a multiple of it,
// with p large enough that the node's extent fits 255 steps and that every
// multiple of 2^p up to the origin plus 255 steps is a float exactly (at most
// 24 significant bits). Returns false when no exponent in range works.
bool axis_grid(double lo, double hi, int& p_out, double& origin_out) {
    const double extent = hi - lo;
    int p = extent > 0.0 ? (int)std::ceil(std::log2(extent / 255.0)) : -126;
    for (; p <= 127; ++p) {
        if (p < -126) continue;
        const double step = std::ldexp(1.0, p);
        const double origin = std::floor(lo / step) * step;
        const double top = origin + 255.0 * step;
        if (top < hi) continue;
        const double mag = std::fmax(std::fabs(origin), std::fabs(top));
        if (mag >= std::ldexp(1.0, p + 24)) continue;  // not exact in float
        p_out = p;
        origin_out = origin;
        return true;
    }
    return false;
}

// Quantized boxes of a wide node's children (always containing the true
// boxes, decoded exactly); false if an axis has no usable grid.
bool quantize(const std::vector<ChildRef>& kids, WideNode& w) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const ChildRef& k : kids)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], (double)k.box.lo[a]);
            hi[a] = std::fmax(hi[a], (double)k.box.hi[a]);
        }
    double origin[3], step[3];
    unsigned exps = (unsigned)kids.size() << 24;
    for (int a = 0; a < 3; ++a) {
        int p;
        if (!axis_grid(lo[a], hi[a], p, origin[a])) return false;
        step[a] = std::ldexp(1.0, p);
        exps |= (unsigned)(p + 127) << (8 * a);
    }
    unsigned qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    for (size_t k = 0; k < kids.size(); ++k)
        for (int a = 0; a < 3; ++a) {
            const double l = std::floor(((double)kids[k].box.lo[a] - origin[a]) / step[a]);
            const double h = std::ceil(((double)kids[k].box.hi[a] - origin[a]) / step[a]);
            if (l < 0.0 || h > 255.0 || l > h) return false;
            qlo[a] |= (unsigned)l << (8 * k);
            qhi[a] |= (unsigned)h << (8 * k);
        }
    w.ox = (float)origin[0];
    w.oy = (float)origin[1];
    w.oz = (float)origin[2];
    w.exps = exps;
    w.lo_x = qlo[0]; w.lo_y = qlo[1]; w.lo_z = qlo[2];
    w.hi_x = qhi[0]; w.hi_y = qhi[1]; w.hi_z = qhi[2];
    w.pad0 = w.pad1 = 0;
    return true;
}

struct Widener {
    const Node* bin;
    int bin_count;
    int leaf_count;
    bool greedy;
    std::vector<WideNode> out;
    int stack_need = 0;
    bool bad = false;

    void open(std::vector<ChildRef>& kids, size_t i) {
        const int c = kids[i].child;
        if (c < 0 || c >= bin_count) {
            bad = true;
            return;
        }
        kids[i] = {bin[c].d.x, node_child_box(bin[c], 0)};
        kids.push_back({bin[c].d.y, node_child_box(bin[c], 1)});
    }

    // Emits the wide node for binary node `node`; `above` counts the stack
    // entries the walk may hold for the ancestors.
    int emit(int node, int above, int depth) {
        if (node < 0 || node >= bin_count || depth > bin_count) {
            bad = true;
            return 0;
        }
        std::vector<ChildRef> kids = {{bin[node].d.x, node_child_box(bin[node], 0)},
                                      {bin[node].d.y, node_child_box(bin[node], 1)}};
        if (greedy) {
            while (kids.size() < 4 && !bad) {
                int best = -1;
                double best_area = -1.0;
                for (size_t i = 0; i < kids.size(); ++i) {
                    if (kids[i].child < 0) continue;
                    const double a = half_area(kids[i].box);
                    if (a > best_area) {
                        best_area = a;
                        best = (int)i;
                    }
                }
                if (best < 0) break;
                open(kids, (size_t)best);
            }
        } else {
            const size_t n0 = kids.size();
            for (size_t i = 0; i < n0 && !bad; ++i)
                if (kids[i].child >= 0) open(kids, i);
        }
        if (bad) return 0;
        const int me = (int)out.size();
        out.push_back(WideNode{});
        int internal = 0;
        for (const ChildRef& k : kids) internal += k.child >= 0;
        const int held = above + (internal > 1 ? internal - 1 : 0);
        if (held > stack_need) stack_need = held;
        int child[4] = {0, 0, 0, 0};
        for (size_t k = 0; k < kids.size(); ++k) {
            const int c = kids[k].child;
            if (c < 0 && ~c >= leaf_count) bad = true;
            child[k] = c < 0 ? c : emit(c, held, depth + 1);
            if (bad) return 0;
        }
        WideNode& w = out[me];  // out may have grown during the recursion
        if (!quantize(kids, w)) {
            bad = true;
            return 0;
        }
        w.child = make_int4(child[0], child[1], child[2], child[3]);
        return me;
    }
};

// Renumbers a wide tree (emitted depth-first, root at 0) so that its first
// kTopNodes nodes in breadth-first order take ids 0 .. kTopNodes - 1, the
// block the walk kernels stage in LDS; every other node keeps its depth-first
// place after them. Child links are remapped; the tree itself is unchanged.
void top_levels_first(std::vector<WideNode>& nodes) {
    const int n = (int)nodes.size();
    std::vector<int> bfs;
    bfs.reserve(kTopNodes);
    bfs.push_back(0);
    for (size_t i = 0; i < bfs.size() && (int)bfs.size() < kTopNodes; ++i) {
        const WideNode& w = nodes[bfs[i]];
        const int count = (int)(w.exps >> 24);
        const int kids[4] = {w.child.x, w.child.y, w.child.z, w.child.w};
        for (int k = 0; k < count && (int)bfs.size() < kTopNodes; ++k)
            if (kids[k] >= 0) bfs.push_back(kids[k]);
    }
    std::vector<int> id(n, -1);
    for (size_t i = 0; i < bfs.size(); ++i) id[bfs[i]] = (int)i;
    int next = (int)bfs.size();
    for (int v = 0; v < n; ++v)
        if (id[v] < 0) id[v] = next++;
    std::vector<WideNode> renumbered(n);
    for (int v = 0; v < n; ++v) {
        WideNode w = nodes[v];
        const int count = (int)(w.exps >> 24);
        int kids[4] = {w.child.x, w.child.y, w.child.z, w.child.w};
        for (int k = 0; k < count; ++k)
            if (kids[k] >= 0) kids[k] = id[kids[k]];
        w.child = make_int4(kids[0], kids[1], kids[2], kids[3]);
        renumbered[id[v]] = w;
    }
    nodes.swap(renumbered);
}

// Collapses a binary tree of `face_count` leaves into wide nodes; returns the
// wide node count or -1 (malformed tree, or no collapse fits the stack).
int widen(const Node* bin, int face_count, std::vector<WideNode>& out) {
    const int bin_count = face_count <= 1 ? 1 : face_count - 1;
    if (face_count == 1) {
        // one triangle: the binary root's second child is an empty placeholder
        if (bin[0].d.x != ~0) return -1;
        WideNode w{};
        const std::vector<ChildRef> only = {{~0, node_child_box(bin[0], 0)}};
        if (!quantize(only, w)) return -1;
        w.child = make_int4(~0, 0, 0, 0);
        out.assign(1, w);
        return 1;
    }
    for (int pass = 0; pass < 2; ++pass) {
        Widener w{bin, bin_count, face_count, pass == 0};
        w.out.reserve(bin_count);
        w.emit(0, 0, 0);
        if (w.bad) return -1;
        if (w.stack_need <= kWideStackSize) {
            out.swap(w.out);
            top_levels_first(out);
            return (int)out.size();
        }
    }
    return -1;
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_bvh_widen(const void* binary_nodes_host, int face_count, void* wide_nodes_out_host,
                            int* wide_count_host) {
    if (face_count <= 0) return wc_fail("wc_bvh_widen: face_count must be positive");
    if (!binary_nodes_host || !wide_nodes_out_host) return wc_fail("wc_bvh_widen: null pointer argument");
    std::vector<WideNode> wide;
    const int count = widen(static_cast<const Node*>(binary_nodes_host), face_count, wide);
    if (count < 0)
        return wc_fail("wc_bvh_widen: the binary tree is malformed or too deep for a %d-entry walk stack",
                       kWideStackSize);
    std::memcpy(wide_nodes_out_host, wide.data(), sizeof(WideNode) * (size_t)count);
    if (wide_count_host) *wide_count_host = count;
    return 0;
}

extern "C" int wc_bvh_build_sah(const float* vertices_host, const int* faces_host, const int* face_group_host,
                                const int* face_material_host, int face_count, void* nodes_out_host,
                                void* sorted_tris_out_host, void* face_tris_out_host) {
    const float* vertices = vertices_host;
    const int* faces = faces_host;
    const int* face_group = face_group_host;
    const int* face_material = face_material_host;
    void* nodes_out = nodes_out_host;
    void* sorted_tris_out = sorted_tris_out_host;
    void* face_tris_out = face_tris_out_host;
    if (face_count < 0) return wc_fail("wc_bvh_build_sah: face_count must be non-negative");
    if (face_count == 0) return 0;
    if (!vertices || !faces || !face_group || !face_material || !nodes_out || !sorted_tris_out || !face_tris_out)
        return wc_fail("wc_bvh_build_sah: null pointer argument");
    const int n = face_count;
    Tri* sorted = static_cast<Tri*>(sorted_tris_out);
    Tri* by_face = static_cast<Tri*>(face_tris_out);
    std::vector<Node> bin_nodes(n <= 1 ? 1 : n - 1);
    Node* nodes = bin_nodes.data();
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
        return wc_bvh_widen(nodes, 1, nodes_out, nullptr);
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
    return wc_bvh_widen(nodes, n, nodes_out, nullptr);
}
