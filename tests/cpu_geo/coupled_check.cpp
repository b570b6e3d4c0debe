// CPU run of the coupled-path device code (wc_coupled.h, built as host C++
// through the shim in hip/): every candidate of every receiver, as the
// coupled kernel evaluates them. Input directory: nodes.bin (64-byte wide
// nodes), tris.bin (48-byte triangles in leaf order), faces.bin (48-byte
// triangles in face order), fgroup.bin / fmat.bin (i32 per face), plane.bin
// (f64 [G,4]), goff.bin (i32 [G+1]), gfaces.bin (i32), mats.bin (f32 [M,5]:
// eps, sigma, mu, gain, thickness), edges.bin (80-byte records), ends.bin
// (f64: tx xyz, tx pol xyz, rx pol xyz, power, frequency), rx.bin (f32 [R,3]).
// Prints one line per receiver: "cell count count_dd re im", then one line
// per found path: "path cell k kind".
#include "hip/hip_runtime.h"
#include "../../witwin-channel_amd/csrc/wc_coupled.h"

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
    const std::string d = argv[1];
    const auto nodes = load<WideNode>(d + "/nodes.bin");
    const auto tris = load<Tri>(d + "/tris.bin");
    const auto faces = load<Tri>(d + "/faces.bin");
    const auto fgroup = load<int>(d + "/fgroup.bin");
    const auto fmat = load<int>(d + "/fmat.bin");
    const auto plane = load<double>(d + "/plane.bin");
    const auto goff = load<int>(d + "/goff.bin");
    const auto gfaces = load<int>(d + "/gfaces.bin");
    const auto mats = load<float>(d + "/mats.bin");
    const auto edges = load<Edge>(d + "/edges.bin");
    const auto ends = load<double>(d + "/ends.bin");
    const auto rx = load<float>(d + "/rx.bin");
    if (nodes.empty() || tris.empty() || goff.empty() || ends.size() < 11 || mats.size() % 5) return 2;
    const int M = (int)mats.size() / 5;
    std::vector<float> eps(M), sig(M), mu(M), gain(M), thick(M);
    for (int m = 0; m < M; ++m) {
        eps[m] = mats[5 * m];
        sig[m] = mats[5 * m + 1];
        mu[m] = mats[5 * m + 2];
        gain[m] = mats[5 * m + 3];
        thick[m] = mats[5 * m + 4];
    }
    SceneView s{};
    s.nodes = nodes.data();
    s.tris = tris.data();
    s.tri_count = (int)tris.size();
    s.node_count = (int)nodes.size();
    s.faces = faces.data();
    s.face_group = fgroup.data();
    s.face_material = fmat.data();
    s.face_count = (int)faces.size();
    s.group_plane = plane.data();
    s.group_face_offset = goff.data();
    s.group_faces = gfaces.data();
    s.group_count = (int)goff.size() - 1;
    s.mat_eps = eps.data();
    s.mat_sigma = sig.data();
    s.mat_mu = mu.data();
    s.mat_gain = gain.data();
    s.mat_thickness = thick.data();
    s.material_count = M;
    s.edges = edges.data();
    s.edge_count = (int)edges.size();
    const V3<double> T{ends[0], ends[1], ends[2]};
    const V3<float> tx_pol{(float)ends[3], (float)ends[4], (float)ends[5]};
    const V3<float> rx_pol{(float)ends[6], (float)ends[7], (float)ends[8]};
    const float sqrt_power = std::sqrt((float)std::max(ends[9], 0.0));
    const double freq = ends[10];
    const double k_d = freq * kTwoPi / kC0;
    const int G = s.group_count, E = s.edge_count;
    const int per_cell = 2 * G * E + E * (E > 0 ? E - 1 : 0);
    for (size_t c = 0; 3 * c + 2 < rx.size(); ++c) {
        const V3<double> X{rx[3 * c], rx[3 * c + 1], rx[3 * c + 2]};
        double re = 0.0, im = 0.0;
        int n = 0, n_dd = 0;
        std::vector<std::pair<int, int>> found;
        for (int k = 0; k < per_cell; ++k) {
            const CoupledRecord r = coupled_candidate(s, k, 7, T, X, tx_pol, rx_pol, sqrt_power, (float)freq, k_d);
            if (r.kind == 0.0f) continue;
            re += r.pf_re;
            im += r.pf_im;
            ++n;
            n_dd += r.kind == 7.0f;
            found.push_back({k, (int)r.kind});
        }
        std::printf("cell %zu %d %d %.9e %.9e\n", c, n, n_dd, re, im);
        for (auto& p : found) std::printf("path %zu %d %d\n", c, p.first, p.second);
    }
    return 0;
}
