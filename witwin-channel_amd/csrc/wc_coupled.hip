// Coupled reflection-diffraction paths of a radiomap (component ids 3
// reflection -> diffraction, 4 diffraction -> reflection, 7 double
// diffraction, summed coherently into one "coupled" slot). One lane evaluates
// one (cell, candidate) with coupled_candidate (wc_coupled.h) and writes its
// path field and gain; a second kernel sums every cell's candidates in
// candidate order in float64 (bit-reproducible).
#include "wc_api_internal.h"
#include "wc_coupled.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

__device__ __forceinline__ int tx_of(const wc_solve& q, int cell) {
    int lo = 0, hi = q.tx_count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.cell_offset[mid] <= cell) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) coupled_eval_kernel(SceneView s, wc_solve q, int cell_begin, int cells,
                                                           int per_cell, int kinds, CoupledRecord* __restrict__ rec) {
    const long long total = (long long)cells * per_cell;
    const double k_d = q.frequency_hz * kTwoPi / kC0;
    const float freq = (float)q.frequency_hz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)blockDim.x * gridDim.x) {
        const int cell = cell_begin + (int)(i / per_cell);
        const int k = (int)(i % per_cell);
        const int t = tx_of(q, cell);
        const int rx_index = q.rx_shared ? cell - q.cell_offset[t] : cell;
        rec[i] = coupled_candidate(s, k, kinds, to_d(ld3(q.tx_pos + 3 * t)), to_d(ld3(q.rx_pos + 3 * rx_index)),
                                   ld3(q.tx_pol + 3 * t), ld3(q.rx_pol + 3 * (q.rx_pol_stride ? rx_index : 0)),
                                   sqrtf(fmaxf(q.tx_power[t], 0.0f)), freq, k_d);
    }
}

// one lane per cell: float64 sums over its candidates in candidate order
__global__ void __launch_bounds__(256) coupled_sum_kernel(int cell_begin, int cells, int per_cell,
                                                          const CoupledRecord* __restrict__ rec,
                                                          float* __restrict__ field, float* __restrict__ power,
                                                          int* __restrict__ count, int* __restrict__ count_dd) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += blockDim.x * gridDim.x) {
        double re = 0.0, im = 0.0, p = 0.0;
        int n = 0, n_dd = 0;
        const CoupledRecord* r = rec + (size_t)c * per_cell;
        for (int k = 0; k < per_cell; ++k) {
            const CoupledRecord x = r[k];
            if (x.kind == 0.0f) continue;
            re += (double)x.pf_re;
            im += (double)x.pf_im;
            p += (double)x.gain;
            ++n;
            n_dd += x.kind == 7.0f ? 1 : 0;
        }
        const int cell = cell_begin + c;
        field[2 * cell] = (float)re;
        field[2 * cell + 1] = (float)im;
        power[cell] = (float)p;
        count[cell] = n;
        if (count_dd) count_dd[cell] = n_dd;
    }
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" long long wc_coupled_candidates_per_cell(int group_count, int edge_count, int kinds) {
    const long long ge = (long long)group_count * edge_count;
    long long n = 0;
    if (kinds & 3) n = 2 * ge;  // R->D and D->R share the layout: both blocks are enumerated
    if (kinds & 4) n = 2 * ge + (long long)edge_count * (edge_count > 0 ? edge_count - 1 : 0);
    return n;
}

extern "C" size_t wc_coupled_workspace_bytes(int cells, int group_count, int edge_count, int kinds) {
    const long long n = wc_coupled_candidates_per_cell(group_count, edge_count, kinds);
    return (size_t)std::max(cells, 0) * (size_t)std::max(n, 1LL) * sizeof(CoupledRecord);
}

extern "C" int wc_coupled_paths(const wc_scene* scene, const wc_solve* solve, int cell_begin, int cell_end, int kinds,
                                const wc_coupled_out* out, void* workspace, size_t workspace_bytes, void* stream) {
    const char* who = "wc_coupled_paths";
    if (!scene || !solve || !out || !out->field || !out->power || !out->count) return wc_fail("%s: null argument", who);
    if (cell_begin < 0 || cell_end < cell_begin || cell_end > solve->total_cells)
        return wc_fail("%s: cell range [%d, %d) outside [0, %d)", who, cell_begin, cell_end, solve->total_cells);
    if ((kinds & ~7) != 0) return wc_fail("%s: kinds must be a combination of 1 (R->D), 2 (D->R), 4 (D->D)", who);
    if (!(solve->frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    const int cells = cell_end - cell_begin;
    const long long per_cell = wc_coupled_candidates_per_cell(s.group_count, s.edge_count, kinds);
    if (cells == 0) return 0;
    if (per_cell > 0x7fffffff) return wc_fail("%s: %lld candidates per cell", who, per_cell);
    if (workspace_bytes < wc_coupled_workspace_bytes(cells, s.group_count, s.edge_count, kinds) || !workspace)
        return wc_fail("%s: workspace too small", who);
    hipStream_t st = (hipStream_t)stream;
    CoupledRecord* rec = static_cast<CoupledRecord*>(workspace);
    if (per_cell > 0) {
        const long long total = (long long)cells * per_cell;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(coupled_eval_kernel, dim3(blocks), dim3(256), 0, st, s, *solve, cell_begin, cells,
                           (int)per_cell, kinds, rec);
    }
    hipLaunchKernelGGL(coupled_sum_kernel, dim3(std::max(1, std::min((cells + 255) / 256, 65536))), dim3(256), 0,
                       st, cell_begin, cells, (int)per_cell, rec, out->field, out->power, out->count,
                       out->count_dd);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}
