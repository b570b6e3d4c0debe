// Row-level field operators (one path row per lane) and the deterministic
// flat accumulator.
#include <mutex>

#include "wc_api_internal.h"
#include "wc_paths.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

__device__ __forceinline__ void store_c(float* p, int64_t i, const Cx<float>& v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}

__device__ __forceinline__ void store_c3(float* p, int64_t i, const C3<float>& v) {
    float* q = p + 6 * i;
    q[0] = v.x.re; q[1] = v.x.im; q[2] = v.y.re; q[3] = v.y.im; q[4] = v.z.re; q[5] = v.z.im;
}

__global__ void free_space_kernel(int count, const float* __restrict__ source, const float* __restrict__ target,
                                  const float* __restrict__ tx_power, const float* __restrict__ tx_pol,
                                  const float* __restrict__ rx_pol, double freq, float* __restrict__ field_vector,
                                  float* __restrict__ coefficient, float* __restrict__ path_field,
                                  float* __restrict__ path_gain, float* __restrict__ path_length,
                                  float* __restrict__ delay, float* __restrict__ direction) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        const V3<float> a = ld3(source + 3 * i), b = ld3(target + 3 * i);
        const V3<double> off = to_d(b) - to_d(a);
        const double dist = norm(off);
        const V3<float> dir = safe_normalize(b - a, v3c<float>(0.f, 0.f, 1.f));
        const double k_d = freq * kTwoPi / kC0;
        const C3<float> v = los_field(dir, ld3(tx_pol + 3 * i), (float)k_d, k_d, dist);
        const Cx<float> coef = project_receiver(v, dir, ld3(rx_pol + 3 * i));
        const Cx<float> pf = scale(coef, sqrtf(fmaxf(tx_power[i], 0.0f)));
        store_c3(field_vector, i, v);
        store_c(coefficient, i, coef);
        store_c(path_field, i, pf);
        path_gain[i] = abs2(pf);
        path_length[i] = (float)dist;
        delay[i] = (float)(dist / kC0);
        direction[3 * i] = dir.x;
        direction[3 * i + 1] = dir.y;
        direction[3 * i + 2] = dir.z;
    }
}

__global__ void project_kernel(int count, const float* __restrict__ field_vector, const float* __restrict__ direction,
                               const float* __restrict__ rx_pol, float* __restrict__ coefficient,
                               float* __restrict__ path_gain) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        const float* f = field_vector + 6 * i;
        const C3<float> v{{f[0], f[1]}, {f[2], f[3]}, {f[4], f[5]}};
        const Cx<float> c = project_receiver(v, ld3(direction + 3 * i), ld3(rx_pol + 3 * i));
        store_c(coefficient, i, c);
        path_gain[i] = abs2(c);
    }
}

__global__ void reflection_sequence_kernel(int count, int depth, const float* __restrict__ source,
                                           const float* __restrict__ target, const float* __restrict__ positions,
                                           const float* __restrict__ normals, const float* __restrict__ tx_power,
                                           const float* __restrict__ tx_pol, const float* __restrict__ rx_pol,
                                           const float* __restrict__ eps_r, const float* __restrict__ sigma_e,
                                           const float* __restrict__ mu_r, const float* __restrict__ gain,
                                           const float* __restrict__ thickness, double freq,
                                           float* __restrict__ field_vector, float* __restrict__ coefficient,
                                           float* __restrict__ path_field, float* __restrict__ path_gain,
                                           float* __restrict__ path_length, float* __restrict__ delay,
                                           float* __restrict__ direction) {
    const V3<float> ez = v3c<float>(0.f, 0.f, 1.f);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        V3<float> prev = ld3(source + 3 * i);
        const V3<float> first = ld3(positions + 3 * (i * depth));
        V3<float> incident = safe_normalize(first - prev, ez);
        C3<float> value = c3_from_real(transverse(incident, ld3(tx_pol + 3 * i)));
        V3<float> outgoing = incident;
        double total = 0.0;
        for (int b = 0; b < depth; ++b) {
            const V3<float> hit = ld3(positions + 3 * (i * depth + b));
            incident = safe_normalize(hit - prev, outgoing);
            total += norm(to_d(hit) - to_d(prev));
            const int m = i * depth + b;
            value = reflect_c3(value, incident, ld3(normals + 3 * m), eps_r[m], sigma_e[m], mu_r[m], gain[m],
                               thickness[m], (float)freq, outgoing);
            prev = hit;
        }
        const V3<float> tgt = ld3(target + 3 * i);
        const V3<float> fdir = safe_normalize(tgt - prev, outgoing);
        total += norm(to_d(tgt) - to_d(prev));
        const double k_d = freq * kTwoPi / kC0;
        const float amp = 1.0f / (2.0f * (float)k_d * fmaxf((float)total, kEpsF));
        value = c3_scale(value, scale(cexpj(neg_kd_phase(k_d, total)), amp));
        const Cx<float> coef = project_receiver(value, fdir, ld3(rx_pol + 3 * i));
        const Cx<float> pf = scale(coef, sqrtf(fmaxf(tx_power[i], 0.0f)));
        store_c3(field_vector, i, value);
        store_c(coefficient, i, coef);
        store_c(path_field, i, pf);
        path_gain[i] = abs2(pf);
        path_length[i] = (float)total;
        delay[i] = (float)(total / kC0);
        direction[3 * i] = fdir.x;
        direction[3 * i + 1] = fdir.y;
        direction[3 * i + 2] = fdir.z;
    }
}

// ---- line-of-sight export of every (transmitter, receiver) pair -------------------
// pair = rx * T + tx; gain = P (lambda / (4 pi d))^2 (|p|^2 - (p . k)^2), d
// clamped at 1e-6 m; no visibility test (reference los.cu:74-125).
__global__ void los_export_kernel(long long pairs, int tx_count, int rx_count, const float* __restrict__ tx_pos,
                                  const float* __restrict__ tx_power, const float* __restrict__ rx_pos,
                                  const float* __restrict__ tx_pol, float scale, int* __restrict__ tx_id,
                                  int* __restrict__ rx_id, float* __restrict__ path_length, float* __restrict__ delay,
                                  float* __restrict__ path_gain, float* __restrict__ gain_matrix) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < pairs;
         i += (long long)blockDim.x * gridDim.x) {
        const int r = (int)(i / tx_count);
        const int t = (int)(i - (long long)r * tx_count);
        const V3<float> a = ld3(tx_pos + 3 * t), b = ld3(rx_pos + 3 * r);
        const V3<float> off = a - b;
        const float d = fmaxf(norm(off), 1.0e-6f);
        const V3<float> k = off * (1.0f / d);
        const V3<float> p = ld3(tx_pol + 3 * t);
        const float pk = dot(p, k);
        const float pattern = fmaxf(dot(p, p) - pk * pk, 0.0f);
        const float g = scale / d;
        const float gain = tx_power[t] * g * g * pattern;
        tx_id[i] = t;
        rx_id[i] = r;
        path_length[i] = d;
        delay[i] = d / (float)kC0;
        path_gain[i] = gain;
        gain_matrix[(size_t)t * rx_count + r] = gain;
    }
}

// ---- deterministic accumulation -------------------------------------------------

__device__ __forceinline__ int accum_slot(int component) {
    if (component >= 0 && component < 3) return component;
    if (component == 3 || component == 4 || component == 7) return 5;
    if (component == 5) return 3;
    if (component == 6) return 4;
    return -1;
}

// key = cell (tx * R + rx) for valid, in-range rows with a slot; invalid rows
// sort after every cell (key = cells).
__global__ void accum_keys_kernel(int count, const unsigned char* __restrict__ valid, const int* __restrict__ tx_id,
                                  const int* __restrict__ rx_id, const int* __restrict__ component, int num_tx,
                                  int num_rx, unsigned int* __restrict__ keys, int* __restrict__ values) {
    const unsigned int cells = (unsigned int)num_tx * (unsigned int)num_rx;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        const int t = tx_id[i], r = rx_id[i];
        const bool ok = valid[i] && accum_slot(component[i]) >= 0 && t >= 0 && r >= 0 && t < num_tx && r < num_rx;
        keys[i] = ok ? (unsigned int)t * (unsigned int)num_rx + (unsigned int)r : cells;
        values[i] = i;
    }
}

__global__ void accum_bounds_kernel(int count, const unsigned int* __restrict__ keys, int cells,
                                    int* __restrict__ start) {
    // start[c] = first sorted position with key >= c (lower bound); start[cells] = first invalid
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= cells; c += blockDim.x * gridDim.x) {
        int lo = 0, hi = count;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (keys[mid] < (unsigned int)c) lo = mid + 1; else hi = mid;
        }
        start[c] = lo;
    }
}

__global__ void accum_reduce_kernel(int cells, const int* __restrict__ start, const int* __restrict__ order,
                                    const int* __restrict__ component, const float* __restrict__ path_gain,
                                    const float* __restrict__ field_real, const float* __restrict__ field_imag,
                                    int coherent, float* __restrict__ component_power,
                                    float* __restrict__ component_field, float* __restrict__ power_total,
                                    float* __restrict__ field_total) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += blockDim.x * gridDim.x) {
        double pw[6] = {0, 0, 0, 0, 0, 0}, re[6] = {0, 0, 0, 0, 0, 0}, im[6] = {0, 0, 0, 0, 0, 0};
        for (int k = start[c]; k < start[c + 1]; ++k) {
            const int row = order[k];
            const int slot = accum_slot(component[row]);
            // float sums in row order mirror the reference's per-slot float accumulation
            pw[slot] += (double)path_gain[row];
            re[slot] += (double)field_real[row];
            im[slot] += (double)field_imag[row];
        }
        double rs = 0.0, is = 0.0, ps = 0.0, scat = 0.0;
        for (int slot = 0; slot < 6; ++slot) {
            const size_t o = (size_t)slot * cells + c;
            double p = pw[slot];
            if (coherent) {
                if (slot == 4) {
                    scat += p;
                } else {
                    p = re[slot] * re[slot] + im[slot] * im[slot];
                }
            } else {
                ps += p;
            }
            component_power[o] = (float)p;
            component_field[2 * o] = (float)re[slot];
            component_field[2 * o + 1] = (float)im[slot];
            if (!coherent || slot != 4) {
                rs += re[slot];
                is += im[slot];
            }
        }
        if (coherent) {
            power_total[c] = (float)(rs * rs + is * is + scat);
            field_total[2 * c] = (float)rs;
            field_total[2 * c + 1] = (float)is;
        } else {
            power_total[c] = (float)ps;
            field_total[2 * c] = (float)sqrt(fmax(ps, 0.0));
            field_total[2 * c + 1] = 0.0f;
        }
    }
}

int blocks_for(long long n, int threads) {
    long long b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 1048576) b = 1048576;
    return (int)b;
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_field_free_space(int count, const float* source, const float* target, const float* tx_power,
                                   const float* tx_polarization, const float* rx_polarization, double frequency_hz,
                                   float* field_vector, float* coefficient, float* path_field, float* path_gain,
                                   float* path_length, float* delay, float* direction, void* stream) {
    if (count < 0) return wc_fail("wc_field_free_space: count must be non-negative");
    if (count == 0) return 0;
    if (!(frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    hipLaunchKernelGGL(free_space_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, count,
                       source, target, tx_power, tx_polarization, rx_polarization, frequency_hz, field_vector,
                       coefficient, path_field, path_gain, path_length, delay, direction);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_field_free_space", err);
}

extern "C" int wc_path_los_export(int tx_count, int rx_count, const float* tx_positions, const float* tx_power,
                                  const float* rx_positions, const float* tx_polarizations, double frequency_hz,
                                  int* tx_id, int* rx_id, float* path_length, float* delay, float* path_gain,
                                  float* path_gain_matrix, void* stream) {
    if (tx_count < 0 || rx_count < 0) return wc_fail("wc_path_los_export: counts must be non-negative");
    if (!(frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    const long long pairs = (long long)tx_count * rx_count;
    if (pairs == 0) return 0;
    const float scale = (float)(kC0 / frequency_hz / (4.0 * kPi));
    hipLaunchKernelGGL(los_export_kernel, dim3(blocks_for(pairs, 256)), dim3(256), 0, (hipStream_t)stream, pairs,
                       tx_count, rx_count, tx_positions, tx_power, rx_positions, tx_polarizations, scale, tx_id, rx_id,
                       path_length, delay, path_gain, path_gain_matrix);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_path_los_export", err);
}

extern "C" int wc_field_project_complex3(int count, const float* field_vector, const float* direction,
                                         const float* rx_polarization, float* coefficient, float* path_gain,
                                         void* stream) {
    if (count < 0) return wc_fail("wc_field_project_complex3: count must be non-negative");
    if (count == 0) return 0;
    hipLaunchKernelGGL(project_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, count,
                       field_vector, direction, rx_polarization, coefficient, path_gain);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_field_project_complex3", err);
}

extern "C" int wc_field_reflection_sequence(int count, int depth, const float* source, const float* target,
                                            const float* interaction_positions, const float* interaction_normals,
                                            const float* tx_power, const float* tx_polarization,
                                            const float* rx_polarization, const float* eps_r, const float* sigma_e,
                                            const float* mu_r, const float* gain, const float* thickness,
                                            double frequency_hz, float* field_vector, float* coefficient,
                                            float* path_field, float* path_gain, float* path_length, float* delay,
                                            float* direction, void* stream) {
    if (count < 0) return wc_fail("wc_field_reflection_sequence: count must be non-negative");
    if (depth < 1) return wc_fail("reflection depth must be positive");
    if (count == 0) return 0;
    if (!(frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    hipLaunchKernelGGL(reflection_sequence_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream,
                       count, depth, source, target, interaction_positions, interaction_normals, tx_power,
                       tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness, frequency_hz,
                       field_vector, coefficient, path_field, path_gain, path_length, delay, direction);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_field_reflection_sequence", err);
}

extern "C" size_t wc_accumulate_workspace_bytes(int count, int num_tx, int num_rx) {
    const size_t n = (size_t)std::max(count, 1);
    const size_t cells = (size_t)std::max(num_tx, 0) * (size_t)std::max(num_rx, 0);
    return align_up(4 * n) * 4 + align_up(4 * radix_sort_hist_ints(count)) + align_up(4 * (cells + 1));
}

namespace wc {
namespace {

// Workspace layout of the flat accumulation: sort keys and row ids (two
// buffers each), the radix histogram, then cells + 1 segment starts. Every
// byte is written before it is read, so nothing needs clearing.
int accumulate_flat_launch(int count, const unsigned char* valid, const int* tx_id, const int* rx_id,
                           const int* component_id, const float* path_gain, const float* field_real,
                           const float* field_imag, int num_tx, int num_rx, int cells, int coherent,
                           float* component_power, float* component_field, float* power_total, float* field_total,
                           void* workspace, hipStream_t st) {
    char* base = (char*)workspace;
    const size_t n = (size_t)std::max(count, 1);
    unsigned int* keys = (unsigned int*)base; base += align_up(4 * n);
    unsigned int* keys_tmp = (unsigned int*)base; base += align_up(4 * n);
    int* vals = (int*)base; base += align_up(4 * n);
    int* vals_tmp = (int*)base; base += align_up(4 * n);
    int* hist = (int*)base; base += align_up(4 * radix_sort_hist_ints(count));
    int* bounds = (int*)base;  // cells + 1 segment starts
    if (count > 0) {
        hipLaunchKernelGGL(accum_keys_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, count, valid, tx_id,
                           rx_id, component_id, num_tx, num_rx, keys, vals);
        int bits = 1;
        while (bits < 32 && ((unsigned long long)cells >> bits) != 0) ++bits;
        if (radix_sort_pairs(keys, vals, keys_tmp, vals_tmp, hist, count, bits, st) != 0)
            return wc_fail("wc_accumulate_flat: radix sort launch failed");
    }
    hipLaunchKernelGGL(accum_bounds_kernel, dim3(blocks_for((long long)cells + 1, 256)), dim3(256), 0, st, count, keys,
                       cells, bounds);
    hipLaunchKernelGGL(accum_reduce_kernel, dim3(blocks_for(cells, 256)), dim3(256), 0, st, cells, bounds, vals,
                       component_id, path_gain, field_real, field_imag, coherent, component_power, component_field,
                       power_total, field_total);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_accumulate_flat", err);
}

// Scratch of the allocation-free accumulation: one copy per GPU, resident
// until the process ends. Host code never uses this name as an address; it
// uses the per-device pointer from hipGetSymbolAddress (scratch_address).
constexpr size_t kAccumulateScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kAccumulateScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; nullptr after a failure, with
// the error recorded.
void* scratch_address() {
    int device = -1;
    hipError_t err = hipGetDevice(&device);
    if (err != hipSuccess) {
        wc_fail_hip("wc_accumulate_flat_noalloc: hipGetDevice", err);
        return nullptr;
    }
    if (device < 0 || device >= kMaxScratchDevices) {
        wc_fail("wc_accumulate_flat_noalloc: device ordinal %d out of range", device);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr) {
            wc_fail_hip("wc_accumulate_flat_noalloc: hipGetSymbolAddress(scratch)", err);
            return nullptr;
        }
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

// 0 and the cell count, or the failure status with the error recorded
int accumulate_sizes(int count, int num_tx, int num_rx, const char* who, int& cells) {
    if (count < 0 || num_tx < 0 || num_rx < 0) return wc_fail("%s: sizes must be non-negative", who);
    const long long cells_ll = (long long)num_tx * num_rx;
    if (cells_ll >= 0x7fffffffLL) return wc_fail("%s: num_tx * num_rx exceeds int32", who);
    cells = (int)cells_ll;
    return 0;
}

}  // namespace
}  // namespace wc

extern "C" int wc_accumulate_flat(int count, const unsigned char* valid, const int* tx_id, const int* rx_id,
                                  const int* component_id, const float* path_gain, const float* field_real,
                                  const float* field_imag, int num_tx, int num_rx, int coherent,
                                  float* component_power, float* component_field, float* power_total,
                                  float* field_total, void* workspace, size_t workspace_bytes, void* stream) {
    int cells = 0;
    if (const int st = accumulate_sizes(count, num_tx, num_rx, "wc_accumulate_flat", cells)) return st;
    if (workspace_bytes < wc_accumulate_workspace_bytes(count, num_tx, num_rx))
        return wc_fail("wc_accumulate_flat: workspace too small");
    if (cells == 0) return 0;
    return accumulate_flat_launch(count, valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, num_tx,
                                  num_rx, cells, coherent, component_power, component_field, power_total, field_total,
                                  workspace, (hipStream_t)stream);
}

extern "C" size_t wc_accumulate_noalloc_max_bytes(void) { return kAccumulateScratchBytes; }

extern "C" int wc_accumulate_flat_noalloc(int count, const unsigned char* valid, const int* tx_id, const int* rx_id,
                                          const int* component_id, const float* path_gain, const float* field_real,
                                          const float* field_imag, int num_tx, int num_rx, int coherent,
                                          float* component_power, float* component_field, float* power_total,
                                          float* field_total, void* stream) {
    int cells = 0;
    if (const int st = accumulate_sizes(count, num_tx, num_rx, "wc_accumulate_flat_noalloc", cells)) return st;
    const size_t need = wc_accumulate_workspace_bytes(count, num_tx, num_rx);
    if (need > kAccumulateScratchBytes)
        return wc_fail("wc_accumulate_flat_noalloc: %d rows over %d cells need %zu bytes of scratch, more than its "
                       "fixed %zu bytes; use wc_accumulate_flat with a workspace", count, cells, need,
                       kAccumulateScratchBytes);
    void* ws = scratch_address();
    if (!ws) return 1;
    if (cells == 0) return 0;
    return accumulate_flat_launch(count, valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, num_tx,
                                  num_rx, cells, coherent, component_power, component_field, power_total, field_total,
                                  ws, (hipStream_t)stream);
}
