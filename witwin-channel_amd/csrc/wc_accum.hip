// The reference's flat accumulation operators with their adjoint and tangent
// companions, in float32 and float64 (reference
// native/channel/kernels/deterministic.cu:829-1262,
// witwin/channel/kernels/deterministic.py:23-390).
//
// Rows scatter (valid, component slot, transmitter, receiver) into six slots
// per cell: 0/1/2 los / reflection / diffraction, 3 transmission, 4
// scattering (a power-domain slot), 5 coupled. Coherent cells square the
// summed field of the five field slots (|S|^2) and add the scattering power;
// incoherent cells sum the powers and expose sqrt(P) as a pseudo-field. The
// arithmetic is the reference's in the output precision: per-slot sums in row
// order (rows are sorted by cell with a stable radix sort, so a cell's rows
// keep their order and the result does not depend on scheduling), |z|^2 as
// fma(re, re, im * im). The backward is one gather per row through the same
// gates with the |.|^2 and sqrt linearized at the forward's cell values; the
// tangent sums the row tangents per cell slot and maps them through the
// linearized finalize.
#include "wc_api_internal.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

constexpr int kSlots = 6;
constexpr int kScatterSlot = 4;

__device__ __forceinline__ int slot_of_row(int component) {
    if (component >= 0 && component < 3) return component;
    if (component == 3 || component == 4 || component == 7) return 5;
    if (component == 5) return 3;
    if (component == 6) return kScatterSlot;
    return -1;
}

__device__ __forceinline__ float sq2(float a, float b) { return fmaf(a, a, b * b); }
__device__ __forceinline__ double sq2(double a, double b) { return fma(a, a, b * b); }
__device__ __forceinline__ float root(float v) { return sqrtf(v); }
__device__ __forceinline__ double root(double v) { return sqrt(v); }

// the cell of a kept row, or -1
__device__ __forceinline__ long long row_cell(const unsigned char* valid, const int* tx_id, const int* rx_id,
                                              const int* component, int i, int num_tx, int num_rx, int& slot) {
    slot = -1;
    if (!valid[i]) return -1;
    const int s = slot_of_row(component[i]);
    const int t = tx_id[i], r = rx_id[i];
    if (s < 0 || t < 0 || r < 0 || t >= num_tx || r >= num_rx) return -1;
    slot = s;
    return (long long)t * num_rx + r;
}

__global__ void acc_keys_kernel(int count, const unsigned char* __restrict__ valid, const int* __restrict__ tx_id,
                                const int* __restrict__ rx_id, const int* __restrict__ component, int num_tx,
                                int num_rx, unsigned* __restrict__ keys, int* __restrict__ vals) {
    const unsigned cells = (unsigned)num_tx * (unsigned)num_rx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        int slot;
        const long long c = row_cell(valid, tx_id, rx_id, component, i, num_tx, num_rx, slot);
        keys[i] = c < 0 ? cells : (unsigned)c;
        vals[i] = i;
    }
}

__global__ void acc_bounds_kernel(int count, const unsigned* __restrict__ keys, int cells, int* __restrict__ start) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= cells; c += blockDim.x * gridDim.x) {
        int lo = 0, hi = count;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (keys[mid] < (unsigned)c) lo = mid + 1;
            else hi = mid;
        }
        start[c] = lo;
    }
}

// per-cell slot sums of the kept rows in row order (T arithmetic), then the
// finalize; `tangent` sums row tangents instead and writes the linearized
// finalize (the forward values come from `fwd_*`)
template <typename T, bool kTangent>
__global__ void acc_cell_kernel(int cells, const int* __restrict__ start, const int* __restrict__ order,
                                const int* __restrict__ component, const T* __restrict__ gain,
                                const T* __restrict__ fre, const T* __restrict__ fim, int coherent,
                                int scattering_coherent, const T* __restrict__ fwd_cfre,
                                const T* __restrict__ fwd_cfim, const T* __restrict__ fwd_power,
                                T* __restrict__ power_total, T* __restrict__ field_total_re,
                                T* __restrict__ field_total_im, T* __restrict__ cpower, T* __restrict__ cfre,
                                T* __restrict__ cfim) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += blockDim.x * gridDim.x) {
        T pw[kSlots], re[kSlots], im[kSlots];
#pragma unroll
        for (int s = 0; s < kSlots; ++s) pw[s] = re[s] = im[s] = T(0);
        for (int k = start[c]; k < start[c + 1]; ++k) {
            const int row = order[k];
            const int s = slot_of_row(component[row]);
            const bool scatter_gain = kTangent ? (s == kScatterSlot ? !scattering_coherent : !coherent) : true;
            if (gain && scatter_gain) pw[s] += gain[row];
            if (fre) re[s] += fre[row];
            if (fim) im[s] += fim[row];
        }
        T rs = T(0), is = T(0), ps = T(0), sc = T(0), frs = T(0), fis = T(0);
        for (int s = 0; s < kSlots; ++s) {
            const size_t o = (size_t)s * cells + c;
            T p = pw[s];
            if (!kTangent) {
                if (coherent) {
                    if (s == kScatterSlot) {
                        if (scattering_coherent) p = sq2(re[s], im[s]);
                        sc += p;
                        cpower[o] = p;
                        cfre[o] = re[s];
                        cfim[o] = im[s];
                        continue;
                    }
                    p = sq2(re[s], im[s]);
                } else {
                    if (s == kScatterSlot && scattering_coherent) p = sq2(re[s], im[s]);
                    ps += p;
                }
            } else {
                const T r0 = fwd_cfre[o], i0 = fwd_cfim[o];
                if (coherent) {
                    if (s == kScatterSlot) {
                        if (scattering_coherent) p = T(2) * (r0 * re[s] + i0 * im[s]);
                        sc += p;
                        cpower[o] = p;
                        cfre[o] = re[s];
                        cfim[o] = im[s];
                        continue;
                    }
                    p = T(2) * (r0 * re[s] + i0 * im[s]);
                    frs += r0;
                    fis += i0;
                } else {
                    if (s == kScatterSlot && scattering_coherent) p = T(2) * (r0 * re[s] + i0 * im[s]);
                    ps += p;
                }
            }
            cpower[o] = p;
            cfre[o] = re[s];
            cfim[o] = im[s];
            rs += re[s];
            is += im[s];
        }
        if (coherent) {
            field_total_re[c] = rs;
            field_total_im[c] = is;
            power_total[c] = kTangent ? T(2) * (frs * rs + fis * is) + sc : sq2(rs, is) + sc;
        } else {
            power_total[c] = ps;
            if (kTangent) {
                const T total = fwd_power[c];
                field_total_re[c] = total > T(0) ? ps / (T(2) * root(total)) : T(0);
            } else {
                field_total_re[c] = root(ps > T(0) ? ps : T(0));
            }
            field_total_im[c] = T(0);
        }
    }
}

// adjoint: one gather per row through the forward's gates
template <typename T>
__global__ void acc_backward_kernel(int count, const unsigned char* __restrict__ valid, const int* __restrict__ tx_id,
                                    const int* __restrict__ rx_id, const int* __restrict__ component, int num_tx,
                                    int num_rx, int coherent, int scattering_coherent, const T* __restrict__ cfre,
                                    const T* __restrict__ cfim, const T* __restrict__ ftre, const T* __restrict__ ftim,
                                    const T* __restrict__ ptot, const T* __restrict__ g_ptot,
                                    const T* __restrict__ g_ftre, const T* __restrict__ g_ftim,
                                    const T* __restrict__ g_cpow, const T* __restrict__ g_cfre,
                                    const T* __restrict__ g_cfim, T* __restrict__ g_gain, T* __restrict__ g_re,
                                    T* __restrict__ g_im) {
    const long long cells = (long long)num_tx * num_rx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        int s;
        const long long cell = row_cell(valid, tx_id, rx_id, component, i, num_tx, num_rx, s);
        T gr = T(0), gi = T(0), gg = T(0);
        if (cell >= 0) {
            const size_t o = (size_t)s * cells + cell;
            if (g_cfre) gr += g_cfre[o];
            if (g_cfim) gi += g_cfim[o];
            if (scattering_coherent && s == kScatterSlot) {
                const T sr = cfre[o], si = cfim[o];
                if (g_cpow) {
                    gr += T(2) * sr * g_cpow[o];
                    gi += T(2) * si * g_cpow[o];
                }
                if (g_ptot) {
                    gr += T(2) * sr * g_ptot[cell];
                    gi += T(2) * si * g_ptot[cell];
                }
                if (!coherent && g_ftre && ptot[cell] > T(0)) {
                    const T f = g_ftre[cell] / (T(2) * root(ptot[cell]));
                    gr += T(2) * sr * f;
                    gi += T(2) * si * f;
                }
            } else if (coherent && s == kScatterSlot) {
                if (g_cpow) gg += g_cpow[o];
                if (g_ptot) gg += g_ptot[cell];
            } else if (coherent) {
                if (g_cpow) {
                    gr += T(2) * cfre[o] * g_cpow[o];
                    gi += T(2) * cfim[o] * g_cpow[o];
                }
                if (g_ftre) gr += g_ftre[cell];
                if (g_ftim) gi += g_ftim[cell];
                if (g_ptot) {
                    gr += T(2) * ftre[cell] * g_ptot[cell];
                    gi += T(2) * ftim[cell] * g_ptot[cell];
                }
            } else {
                if (g_cpow) gg += g_cpow[o];
                if (g_ptot) gg += g_ptot[cell];
                if (g_ftre && ptot[cell] > T(0)) gg += g_ftre[cell] / (T(2) * root(ptot[cell]));
            }
        }
        g_gain[i] = gg;
        g_re[i] = gr;
        g_im[i] = gi;
    }
}

size_t align_ws(size_t n) { return (n + 255) & ~(size_t)255; }

struct AccLayout {
    size_t keys, keys_tmp, vals, vals_tmp, hist, bounds, total;
};

AccLayout acc_layout(int count, long long cells) {
    const size_t n = (size_t)std::max(count, 1);
    AccLayout l{};
    size_t o = 0;
    l.keys = o;
    o += align_ws(4 * n);
    l.keys_tmp = o;
    o += align_ws(4 * n);
    l.vals = o;
    o += align_ws(4 * n);
    l.vals_tmp = o;
    o += align_ws(4 * n);
    l.hist = o;
    o += align_ws(4 * radix_sort_hist_ints(count));
    l.bounds = o;
    o += align_ws(4 * (size_t)(cells + 1));
    l.total = o;
    return l;
}

int grid_of(long long n) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

int check_sizes(int count, int num_tx, int num_rx, const char* who) {
    if (count < 0 || num_tx < 0 || num_rx < 0) return wc_fail("%s: sizes must be non-negative", who);
    if ((long long)num_tx * num_rx >= 0xffffffffLL) return wc_fail("%s: too many cells", who);
    return 0;
}

// sorted row order by cell and per-cell bounds in the workspace
int sort_rows(int count, const unsigned char* valid, const int* tx_id, const int* rx_id, const int* component,
              int num_tx, int num_rx, void* workspace, hipStream_t st, const int** order, const int** bounds) {
    const long long cells = (long long)num_tx * num_rx;
    const AccLayout l = acc_layout(count, cells);
    char* ws = static_cast<char*>(workspace);
    unsigned* keys = reinterpret_cast<unsigned*>(ws + l.keys);
    int* vals = reinterpret_cast<int*>(ws + l.vals);
    int* start = reinterpret_cast<int*>(ws + l.bounds);
    if (count > 0) {
        hipLaunchKernelGGL(acc_keys_kernel, dim3(grid_of(count)), dim3(256), 0, st, count, valid, tx_id, rx_id,
                           component, num_tx, num_rx, keys, vals);
        int bits = 1;
        while (bits < 32 && ((unsigned long long)cells >> bits) != 0) ++bits;
        if (radix_sort_pairs(keys, vals, reinterpret_cast<unsigned*>(ws + l.keys_tmp),
                             reinterpret_cast<int*>(ws + l.vals_tmp), reinterpret_cast<int*>(ws + l.hist), count,
                             bits, st) != 0)
            return -1;
    }
    hipLaunchKernelGGL(acc_bounds_kernel, dim3(grid_of(cells + 1)), dim3(256), 0, st, count, keys, (int)cells, start);
    *order = vals;
    *bounds = start;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T>
int forward_t(const wc_accum_rows* r, const wc_accum_out* o, int tangent, const wc_accum_out* fwd, void* workspace,
              hipStream_t st, const char* who) {
    const int* order = nullptr;
    const int* bounds = nullptr;
    if (sort_rows(r->count, r->valid, r->tx_id, r->rx_id, r->component_id, r->num_tx, r->num_rx, workspace, st,
                  &order, &bounds) != 0)
        return wc_fail_hip(who, hipGetLastError());
    const int cells = r->num_tx * r->num_rx;
    if (cells == 0) return 0;
    const T* gain = static_cast<const T*>(r->path_gain);
    const T* fre = static_cast<const T*>(r->field_real);
    const T* fim = static_cast<const T*>(r->field_imag);
    T* pt = static_cast<T*>(o->power_total);
    T* ftr = static_cast<T*>(o->field_total_real);
    T* fti = static_cast<T*>(o->field_total_imag);
    T* cp = static_cast<T*>(o->component_power);
    T* cfr = static_cast<T*>(o->component_field_real);
    T* cfi = static_cast<T*>(o->component_field_imag);
    if (tangent)
        hipLaunchKernelGGL((acc_cell_kernel<T, true>), dim3(grid_of(cells)), dim3(256), 0, st, cells, bounds, order,
                           r->component_id, gain, fre, fim, r->coherent, r->scattering_coherent,
                           static_cast<const T*>(fwd->component_field_real),
                           static_cast<const T*>(fwd->component_field_imag),
                           static_cast<const T*>(fwd->power_total), pt, ftr, fti, cp, cfr, cfi);
    else
        hipLaunchKernelGGL((acc_cell_kernel<T, false>), dim3(grid_of(cells)), dim3(256), 0, st, cells, bounds, order,
                           r->component_id, gain, fre, fim, r->coherent, r->scattering_coherent, (const T*)nullptr,
                           (const T*)nullptr, (const T*)nullptr, pt, ftr, fti, cp, cfr, cfi);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

int check_rows_in(const wc_accum_rows* r, const char* who, bool need_values) {
    if (!r) return wc_fail("%s: null rows", who);
    if (int rc = check_sizes(r->count, r->num_tx, r->num_rx, who)) return rc;
    if (r->dtype != 0 && r->dtype != 1) return wc_fail("%s: dtype must be 0 (float32) or 1 (float64)", who);
    if (r->scattering_coherent != 0 && r->scattering_coherent != 1)
        return wc_fail("scattering_combine_domain must be 0 (power) or 1 (coherent)");
    if (r->count > 0 && (!r->valid || !r->tx_id || !r->rx_id || !r->component_id))
        return wc_fail("%s: row gates are required", who);
    if (need_values && r->count > 0 && (!r->path_gain || !r->field_real || !r->field_imag))
        return wc_fail("%s: path_gain and the field rows are required", who);
    return 0;
}

int check_out(const wc_accum_out* o, const char* who) {
    if (!o || !o->power_total || !o->field_total_real || !o->field_total_imag || !o->component_power ||
        !o->component_field_real || !o->component_field_imag)
        return wc_fail("%s: every output is required", who);
    return 0;
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" size_t wc_accumulate_rows_workspace_bytes(int count, int num_tx, int num_rx) {
    return acc_layout(std::max(count, 0), (long long)std::max(num_tx, 0) * std::max(num_rx, 0)).total;
}

extern "C" int wc_accumulate_rows(const wc_accum_rows* rows, const wc_accum_out* out, void* workspace, void* stream) {
    const char* who = "wc_accumulate_rows";
    if (int rc = check_rows_in(rows, who, true)) return rc;
    if (int rc = check_out(out, who)) return rc;
    if (!workspace) return wc_fail("%s: workspace is required", who);
    return rows->dtype == 1 ? forward_t<double>(rows, out, 0, nullptr, workspace, (hipStream_t)stream, who)
                            : forward_t<float>(rows, out, 0, nullptr, workspace, (hipStream_t)stream, who);
}

extern "C" int wc_accumulate_rows_jvp(const wc_accum_rows* tangents, const wc_accum_out* forward,
                                      const wc_accum_out* out, void* workspace, void* stream) {
    const char* who = "wc_accumulate_rows_jvp";
    if (int rc = check_rows_in(tangents, who, false)) return rc;
    if (int rc = check_out(out, who)) return rc;
    if (!forward || !forward->power_total || !forward->component_field_real || !forward->component_field_imag)
        return wc_fail("%s: the forward power and component fields are required", who);
    if (!workspace) return wc_fail("%s: workspace is required", who);
    return tangents->dtype == 1
               ? forward_t<double>(tangents, out, 1, forward, workspace, (hipStream_t)stream, who)
               : forward_t<float>(tangents, out, 1, forward, workspace, (hipStream_t)stream, who);
}

extern "C" int wc_accumulate_rows_backward(const wc_accum_rows* rows, const wc_accum_out* forward,
                                           const wc_accum_out* grad_out, void* grad_path_gain, void* grad_field_real,
                                           void* grad_field_imag, void* stream) {
    const char* who = "wc_accumulate_rows_backward";
    if (int rc = check_rows_in(rows, who, false)) return rc;
    if (!forward || !forward->power_total || !forward->field_total_real || !forward->field_total_imag ||
        !forward->component_field_real || !forward->component_field_imag)
        return wc_fail("%s: the forward outputs are required", who);
    if (!grad_out) return wc_fail("%s: null cotangents", who);
    if (rows->count == 0) return 0;
    if (!grad_path_gain || !grad_field_real || !grad_field_imag) return wc_fail("%s: gradient outputs required", who);
    hipStream_t st = (hipStream_t)stream;
#define WC_ACC_BACKWARD(T)                                                                                          \
    hipLaunchKernelGGL(acc_backward_kernel<T>, dim3(grid_of(rows->count)), dim3(256), 0, st, rows->count,          \
                       rows->valid, rows->tx_id, rows->rx_id, rows->component_id, rows->num_tx, rows->num_rx,      \
                       rows->coherent, rows->scattering_coherent, (const T*)forward->component_field_real,         \
                       (const T*)forward->component_field_imag, (const T*)forward->field_total_real,                \
                       (const T*)forward->field_total_imag, (const T*)forward->power_total,                         \
                       (const T*)grad_out->power_total, (const T*)grad_out->field_total_real,                       \
                       (const T*)grad_out->field_total_imag, (const T*)grad_out->component_power,                   \
                       (const T*)grad_out->component_field_real, (const T*)grad_out->component_field_imag,          \
                       (T*)grad_path_gain, (T*)grad_field_real, (T*)grad_field_imag)
    if (rows->dtype == 1) WC_ACC_BACKWARD(double);
    else WC_ACC_BACKWARD(float);
#undef WC_ACC_BACKWARD
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}
