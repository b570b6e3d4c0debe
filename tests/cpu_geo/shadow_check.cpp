// CPU check of the transmitter-leg shadow bins and their float pre-test, on
// the device code itself (built as host C++ through the shim in hip/):
//  - every bin a triangle sets is blocked by that triangle for 11 legs sampled
//    across the bin, under the float watertight test with the end bias the
//    leg test uses (so the BVH walk, which reaches every triangle that test
//    hits, blocks the leg too);
//  - the float pre-test never skips a triangle that sets a bin.
// Input directory: tris.bin (Tri rows, group id in v0.w), edges.bin (Edge
// records), txs.bin (float xyz per transmitter). Prints one summary line;
// exit status 1 on any violation.
#include "hip/hip_runtime.h"
#include "../../witwin-channel_amd/csrc/wc_radiomap.hip"

#include <cstdio>
#include <string>
#include <vector>

int wc_fail(const char*, ...) { return 1; }
int wc_fail_hip(const char*, hipError_t) { return 2; }

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
    const auto tris = load<Tri>(dir + "/tris.bin");
    const auto edges = load<Edge>(dir + "/edges.bin");
    const auto txs = load<float>(dir + "/txs.bin");
    if (tris.empty() || edges.empty() || txs.size() < 3) return 2;
    long pairs = 0, set_bins = 0, samples = 0, violations = 0, tests = 0, skipped = 0, skipped_with_bins = 0;
    for (size_t ti = 0; ti + 2 < txs.size(); ti += 3) {
        const V3<float> tx{txs[ti], txs[ti + 1], txs[ti + 2]};
        for (const Edge& e : edges) {
            ++pairs;
            const ShadowFrame f = shadow_frame(e, tx);
            if (!f.valid) continue;
            const FanPlane plane = fan_plane(f);
            for (const Tri& tri : tris) {
                const int group = __float_as_int(tri.v0.w);
                if (group == e.group0 || group == e.group1) continue;
                const unsigned long long bits = triangle_shadow_bins(f, tri);
                ++tests;
                if (fan_plane_misses(plane, tri)) {
                    ++skipped;
                    if (bits) ++skipped_with_bins;
                }
                for (int b = 0; b < 64; ++b) {
                    if (!((bits >> b) & 1ull)) continue;
                    ++set_bins;
                    for (int k = 0; k <= 10; ++k) {
                        const double tc = (b + k / 10.0) * (double)e.length / 64.0;
                        const V3<float> q{e.ax + e.dx * (float)tc, e.ay + e.dy * (float)tc, e.az + e.dz * (float)tc};
                        const V3<float> seg = q - tx;
                        const float len = norm(seg);
                        ++samples;
                        if (!(len > 2.0f * kSegmentBias)) {
                            ++violations;
                            continue;
                        }
                        const float eps = kSegmentBias / len;
                        const RayPre r = ray_setup(tx.x, tx.y, tx.z, seg.x, seg.y, seg.z);
                        if (!(tri_hit(r, tri, eps, 1.0f - eps) >= 0.0f)) ++violations;
                    }
                }
            }
        }
    }
    std::printf("pairs %ld set_bins %ld samples %ld violations %ld tests %ld skipped %ld skipped_with_bins %ld\n",
                pairs, set_bins, samples, violations, tests, skipped, skipped_with_bins);
    return violations || skipped_with_bins ? 1 : 0;
}
