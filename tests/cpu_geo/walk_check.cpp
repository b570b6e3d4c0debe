// CPU check of the 4-wide tree walk (wc_trace.h, built as host C++ through
// the shim in hip/) against a test of every triangle: for each segment, the
// any-hit walk must agree with "some triangle is hit", and the closest-hit
// walk must return the smallest hit distance with the lowest face id among
// equal distances, except where another surface meets the segment within
// the triangle test's rounding of the distance (a few ulps of the vertex
// coordinates over the segment length): there the walk's hit must lie in that
// window. Input directory: nodes.bin (64-byte wide nodes), tris.bin
// (triangles in leaf order), segs.bin (float a xyz, b xyz per segment).
// Prints one summary line; exit status 1 on any disagreement.
#include "hip/hip_runtime.h"
#include "../../witwin-channel_amd/csrc/wc_trace.h"

#include <cstdio>
#include <string>
#include <vector>

using namespace wc;

template <class T> static std::vector<T> load(const std::string& path) {
    std::vector<T> v;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return v;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)n / sizeof(T));
    if (std::fread(v.data(), 1, (size_t)n, f) != (size_t)n) v.clear();
    std::fclose(f);
    return v;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const auto nodes = load<WideNode>(dir + "/nodes.bin");
    const auto tris = load<Tri>(dir + "/tris.bin");
    const auto segs = load<float>(dir + "/segs.bin");
    if (nodes.empty() || tris.empty() || segs.size() < 6) return 2;
    SceneView s{};
    s.nodes = nodes.data();
    s.tris = tris.data();
    s.tri_count = (int)tris.size();
    s.node_count = (int)nodes.size();
    // rounding window of the triangle distance: 64 ulps of the largest
    // coordinate, in units of the segment length
    float coord = 0.0f;
    for (const Tri& t : tris)
        for (const float c : {t.v0.x, t.v0.y, t.v0.z, t.v1.x, t.v1.y, t.v1.z, t.v2.x, t.v2.y, t.v2.z})
            coord = fmaxf(coord, fabsf(c));
    long count = 0, hits = 0, any_wrong = 0, closest_wrong = 0, ties = 0, near_ties = 0;
    for (size_t i = 0; i + 5 < segs.size(); i += 6) {
        const V3<float> a{segs[i], segs[i + 1], segs[i + 2]}, b{segs[i + 3], segs[i + 4], segs[i + 5]};
        const V3<float> d = b - a;
        const float len = norm(d);
        if (!(len > 2.0e-4f)) continue;
        const float eps = 1.0e-4f / len;
        const RayPre r = ray_setup(a.x, a.y, a.z, d.x, d.y, d.z);
        bool brute_any = false;
        float best_t = 1.0f - eps;
        int best_face = -1, at_best = 0;
        for (const Tri& t : tris) {
            const float th = tri_hit(r, t, eps, 1.0f - eps);
            if (th < 0.0f) continue;
            brute_any = true;
            const int f = __float_as_int(t.v1.w);
            if (best_face < 0 || th < best_t) {
                best_t = th;
                best_face = f;
                at_best = 1;
            } else if (th == best_t) {
                ++at_best;
                if (f < best_face) best_face = f;
            }
        }
        ++count;
        hits += brute_any;
        ties += at_best > 1;
        if (occluded(s, r, eps, 1.0f - eps, -1, -1) != brute_any) ++any_wrong;
        const Hit h = closest_hit(s, r, eps, 1.0f - eps, -1);
        if (!brute_any) {
            if (h.face >= 0) ++closest_wrong;
        } else if (!(h.t == best_t && h.face == best_face)) {
            const float window = 64.0f * 1.1920929e-7f * coord / len;
            bool in_window = h.face >= 0 && h.t <= best_t + window;
            if (in_window) {
                // the walk's face must be a real hit at the distance it reports
                bool real = false;
                for (const Tri& t : tris)
                    if (__float_as_int(t.v1.w) == h.face) real = tri_hit(r, t, eps, 1.0f - eps) == h.t;
                in_window = real;
            }
            if (in_window) ++near_ties;
            else ++closest_wrong;
        }
    }
    std::printf("segments %ld hit %ld ties %ld near_ties %ld any_wrong %ld closest_wrong %ld\n", count, hits, ties,
                near_ties, any_wrong, closest_wrong);
    return any_wrong || closest_wrong ? 1 : 0;
}
