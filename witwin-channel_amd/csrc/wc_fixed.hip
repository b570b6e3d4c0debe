// Fixed-topology replay of frozen path rows (consumer reevaluate).
//
// Every row names a component (line of sight or a reflection chain), its
// depth and its frozen face sequence, plus the current source and target. A
// line-of-sight row re-tests its segment; a reflection row re-solves the image
// method over the planes of its faces' coplanar groups (the winners stay
// fixed, the geometry moves with the endpoints), re-checks that every hit
// lands inside a face of its group and that every leg is clear, and recomputes
// the field. These are the same device functions the radiomap and the row
// export use, so a replay at the discovery endpoints reproduces the exported
// rows. A row that no longer exists publishes valid = 0 and exact zeros.
#include "wc_api_internal.h"
#include "wc_paths.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

struct ReplayOut {
    bool valid;
    C3<float> field;
    V3<float> arrival;
    double length_d;
};

template <int D>
__device__ inline ReplayOut replay_reflection(const SceneView& s, const int* frozen_faces, const V3<float>& tx,
                                              const V3<float>& rx, const V3<float>& tx_pol, float freq, double k_d,
                                              V3<float>* hits_out, V3<float>* normals_out) {
    int groups[kMaxDepth];
    double images[kMaxDepth * 3];
    V3<double> p = to_d(tx);
#pragma unroll
    for (int j = 0; j < D; ++j) {
        groups[j] = s.face_group[frozen_faces[j]];
        double off;
        const V3<double> n = group_normal(s, groups[j], off);
        p = p - n * (2.0 * (dot(p, n) - off));
        images[3 * j] = p.x;
        images[3 * j + 1] = p.y;
        images[3 * j + 2] = p.z;
    }
    ReplayOut out;
    out.length_d = norm(to_d(rx) - p);
    V3<double> hd[D];
    int faces[D];
    out.valid = solve_chain<D, true>(s, groups, images, tx, rx, hd, faces);
    V3<float> hf[D], nf[D];
    int mats[D];
#pragma unroll
    for (int b = 0; b < D; ++b) {
        double off;
        nf[b] = to_f(group_normal(s, groups[b], off));
        hf[b] = out.valid ? to_f(hd[b]) : V3<float>{0.0f, 0.0f, 0.0f};
        // materials of the frozen labels (a hit that slid onto a coplanar twin keeps its label)
        mats[b] = s.face_material[frozen_faces[b]];
        hits_out[b] = hf[b];
        normals_out[b] = nf[b];
    }
    out.field = C3<float>{};
    out.arrival = safe_normalize(rx - tx, v3c<float>(0.f, 0.f, 1.f));
    if (out.valid) {
        V3<float> departure;
        out.field = reflection_field<D>(s, tx, rx, hf, nf, mats, tx_pol, freq, k_d, out.length_d, out.arrival,
                                        departure);
    }
    return out;
}

__global__ void __launch_bounds__(256) fixed_rows_kernel(SceneView s, wc_fixed_rows r, double freq_d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.count) return;
    const V3<float> tx = ld3(r.source + 3 * i), rx = ld3(r.target + 3 * i);
    const V3<float> tx_pol = ld3(r.tx_polarization + 3 * i), rx_pol = ld3(r.rx_polarization + 3 * i);
    const double k_d = freq_d * kTwoPi / kC0;
    const float freq = (float)freq_d;
    const int comp = r.component[i];
    const int depth = r.depth[i];
    const int w = r.width;
    V3<float> hits[kMaxDepth], normals[kMaxDepth];
    ReplayOut out;
    out.valid = false;
    out.field = C3<float>{};
    out.arrival = safe_normalize(rx - tx, v3c<float>(0.f, 0.f, 1.f));
    out.length_d = norm(to_d(rx) - to_d(tx));
    if (comp == kCidLos && depth == 0) {
        const V3<float> seg = rx - tx;
        const float len = (float)out.length_d;
        bool clear = true;
        if (len > 2.0f * kSegmentBias) {
            const float eps = kSegmentBias / len;
            const RayPre ray = ray_setup(tx.x, tx.y, tx.z, seg.x, seg.y, seg.z);
            clear = !occluded(s, ray, eps, 1.0f - eps, -1, -1);
        }
        out.valid = clear;
        if (clear) out.field = los_field(out.arrival, tx_pol, (float)k_d, k_d, out.length_d);
    } else if (comp == kCidRefl && depth >= 1 && depth <= kMaxDepth && depth <= w) {
        const int* f = r.faces + (size_t)i * w;
        switch (depth) {
            case 1: out = replay_reflection<1>(s, f, tx, rx, tx_pol, freq, k_d, hits, normals); break;
            case 2: out = replay_reflection<2>(s, f, tx, rx, tx_pol, freq, k_d, hits, normals); break;
            case 3: out = replay_reflection<3>(s, f, tx, rx, tx_pol, freq, k_d, hits, normals); break;
            case 4: out = replay_reflection<4>(s, f, tx, rx, tx_pol, freq, k_d, hits, normals); break;
            default: out = replay_reflection<5>(s, f, tx, rx, tx_pol, freq, k_d, hits, normals); break;
        }
    }
    PathRecord rec;
    finish_record(rec, out.field, out.arrival, rx_pol, sqrtf(fmaxf(r.tx_power[i], 0.0f)));
    r.valid[i] = out.valid ? 1 : 0;
    r.path_field[2 * i] = out.valid ? rec.path_field.re : 0.0f;
    r.path_field[2 * i + 1] = out.valid ? rec.path_field.im : 0.0f;
    const Cx<float> comps[3] = {out.field.x, out.field.y, out.field.z};
    for (int j = 0; j < 3; ++j) {
        r.field_xyz[6 * i + 2 * j] = out.valid ? comps[j].re : 0.0f;
        r.field_xyz[6 * i + 2 * j + 1] = out.valid ? comps[j].im : 0.0f;
    }
    r.path_length[i] = (float)out.length_d;
    r.delay[i] = (float)(out.length_d / kC0);
    r.direction[3 * i] = out.arrival.x;
    r.direction[3 * i + 1] = out.arrival.y;
    r.direction[3 * i + 2] = out.arrival.z;
    for (int b = 0; b < w; ++b) {
        const bool on = comp == kCidRefl && b < depth && depth <= kMaxDepth;
        const V3<float> h = on ? hits[b] : V3<float>{0.0f, 0.0f, 0.0f};
        const V3<float> n = on ? normals[b] : V3<float>{0.0f, 0.0f, 0.0f};
        float* pp = r.positions + ((size_t)i * w + b) * 3;
        float* nn = r.normals + ((size_t)i * w + b) * 3;
        pp[0] = h.x; pp[1] = h.y; pp[2] = h.z;
        nn[0] = n.x; nn[1] = n.y; nn[2] = n.z;
    }
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_fixed_rows_evaluate(const wc_scene* scene, const wc_fixed_rows* rows, double frequency_hz,
                                      void* stream) {
    if (!scene || !rows) return wc_fail("wc_fixed_rows_evaluate: null argument");
    if (rows->count < 0 || rows->width < 0 || rows->width > kMaxDepth)
        return wc_fail("wc_fixed_rows_evaluate: invalid sizes (count %d, width %d)", rows->count, rows->width);
    if (rows->count == 0) return 0;
    if (!(frequency_hz > 0.0)) return wc_fail("wc_fixed_rows_evaluate: frequency_hz must be positive");
    if (!rows->component || !rows->depth || !rows->source || !rows->target || !rows->tx_power ||
        !rows->tx_polarization || !rows->rx_polarization || !rows->valid || !rows->path_field ||
        !rows->field_xyz || !rows->path_length || !rows->delay || !rows->direction)
        return wc_fail("wc_fixed_rows_evaluate: null row pointer");
    if (rows->width > 0 && (!rows->faces || !rows->positions || !rows->normals))
        return wc_fail("wc_fixed_rows_evaluate: face sequence and interaction outputs are required");
    if (scene->face_count > 0 && (!scene->nodes || !scene->tris || !scene->faces))
        return wc_fail("wc_fixed_rows_evaluate: scene tables are null");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(fixed_rows_kernel, dim3((rows->count + 255) / 256), dim3(256), 0, (hipStream_t)stream, s,
                       *rows, frequency_hz);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_fixed_rows_evaluate", err);
}
