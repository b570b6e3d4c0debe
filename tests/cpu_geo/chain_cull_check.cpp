// CPU check of the tile cull of reflection chains (chain_reaches_box) on the
// device code itself: for every 8 x 8 tile of a receiver grid and every
// chain, a chain the cull drops must not back-project into faces of all its
// groups (solve_chain without the leg tests) for any receiver of the tile.
// Input directory: counts.bin [G, F, K, side], plane.bin (G x 4 float64),
// goff.bin / gfaces.bin (group CSR), tris.bin (Tri rows by face), chains.bin
// (K x 5 int32, -1 padded), images.bin (K x 5 x 3 float64), rx.bin (side^2
// float xyz), tx.bin (float xyz). Prints one summary line; exit status 1
// when a dropped chain was valid.
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

template <int D>
static bool valid_somewhere(const SceneView& s, const int* groups, const double* images, const V3<float>& tx,
                            const std::vector<V3<float>>& cells) {
    for (const V3<float>& r : cells) {
        V3<double> h[D];
        int f[D];
        if (solve_chain<D, false>(s, groups, images, tx, r, h, f)) return true;
    }
    return false;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const auto counts = load<int>(dir + "/counts.bin");
    const auto plane = load<double>(dir + "/plane.bin");
    const auto goff = load<int>(dir + "/goff.bin");
    const auto gfaces = load<int>(dir + "/gfaces.bin");
    const auto tris = load<Tri>(dir + "/tris.bin");
    const auto chains = load<int>(dir + "/chains.bin");
    const auto images = load<double>(dir + "/images.bin");
    const auto rx = load<float>(dir + "/rx.bin");
    const auto txv = load<float>(dir + "/tx.bin");
    if (counts.size() < 4 || txv.size() < 3) return 2;
    const int G = counts[0], F = counts[1], K = counts[2], side = counts[3];
    if ((int)plane.size() != 4 * G || (int)tris.size() != F || (int)chains.size() != 5 * K ||
        (int)images.size() != 15 * K || (int)rx.size() != 3 * side * side)
        return 2;
    SceneView s{};
    s.group_plane = plane.data();
    s.group_face_offset = goff.data();
    s.group_faces = gfaces.data();
    s.group_count = G;
    s.faces = tris.data();
    s.face_count = F;
    wc_solve q{};
    q.chain_groups = chains.data();
    q.chain_images = images.data();
    const V3<float> tx{txv[0], txv[1], txv[2]};
    long culled = 0, kept = 0, kept_valid = 0, wrong = 0;
    for (int bi = 0; bi < side; bi += 8)
        for (int bj = 0; bj < side; bj += 8) {
            std::vector<V3<float>> cells;
            V3<float> lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
            for (int i = bi; i < std::min(bi + 8, side); ++i)
                for (int j = bj; j < std::min(bj + 8, side); ++j) {
                    const int r = i * side + j;
                    const V3<float> c{rx[3 * r], rx[3 * r + 1], rx[3 * r + 2]};
                    cells.push_back(c);
                    lo = {fminf(lo.x, c.x), fminf(lo.y, c.y), fminf(lo.z, c.z)};
                    hi = {fmaxf(hi.x, c.x), fmaxf(hi.y, c.y), fmaxf(hi.z, c.z)};
                }
            for (int c = 0; c < K; ++c) {
                const int depth = chains[5 * c + 1] >= 0 ? 2 : 1;
                const bool reach = chain_reaches_box(s, q, c, depth, lo, hi);
                const bool any = depth == 1 ? valid_somewhere<1>(s, &chains[5 * c], &images[15 * c], tx, cells)
                                            : valid_somewhere<2>(s, &chains[5 * c], &images[15 * c], tx, cells);
                if (reach) {
                    ++kept;
                    kept_valid += any;
                } else {
                    ++culled;
                    wrong += any;
                }
            }
        }
    std::printf("culled %ld kept %ld kept_valid %ld wrong %ld\n", culled, kept, kept_valid, wrong);
    return wrong ? 1 : 0;
}
