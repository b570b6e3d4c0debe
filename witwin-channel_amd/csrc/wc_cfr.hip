// Channel frequency response of padded per-pair path sets.
//
//   out[b, s, k] = sum_p valid[b,p] a[b,p,s] exp(-j 2 pi f_k (tau[b,p] - tau0[b]))
//
// One wavefront per (pair b, time step s, block of 64 * kPerLane subcarriers).
// Lane l owns subcarriers k0 + l + 64 i (i < kPerLane), so every store of the
// wave is one contiguous 512-byte run. The phase argument f * tau is reduced
// modulo one cycle in float64 before the float32 sincos, which keeps the phase
// error near 1e-7 rad at the ~1e4-cycle arguments of km paths at mmWave. On a
// uniform grid f_k = f0 + k df each lane seeds its phasor in float64 every
// kReseed subcarriers and advances it by the 64 df step in float32 in between
// (at most kReseed - 1 rotations, below 1e-6 rad of drift); an arbitrary
// frequency list evaluates every phasor directly. The path loop is
// wave-uniform (tau, valid: scalar loads) and the per-lane sums run in
// float32 over paths in index order, like the complex64 reduction they
// replace.
#include "wc_api_internal.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

constexpr int kPerLane = 16;
constexpr int kReseed = 8;             // float64 phasor seed every kReseed subcarriers of a lane
constexpr int kBlock = 64 * kPerLane;  // subcarriers per wavefront
constexpr double kTwoPiD = 6.283185307179586476925286766559;

__device__ __forceinline__ void phasor(double cycles, float& c, float& s) {
    const double frac = cycles - floor(cycles);
    // exp(-j 2 pi frac)
    __sincosf(-(float)(kTwoPiD * frac), &s, &c);
}

__global__ void __launch_bounds__(256) cfr_kernel(const float* __restrict__ a, const float* __restrict__ tau,
                                                  const unsigned char* __restrict__ valid, int pairs, int max_paths,
                                                  int time_steps, const double* __restrict__ frequencies,
                                                  double f0, double df, int freq_count, int blocks_per_row,
                                                  int normalize, float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long long rows = (long long)pairs * time_steps;
    if (wave >= rows * blocks_per_row) return;
    const long long row = wave / blocks_per_row;          // b * S + s
    const int k0 = (int)(wave - row * blocks_per_row) * kBlock;
    const int b = (int)(row / time_steps);
    const int s = (int)(row - (long long)b * time_steps);
    const float* tau_b = tau + (size_t)b * max_paths;
    const unsigned char* valid_b = valid + (size_t)b * max_paths;
    double tau0 = 0.0;
    if (normalize) {
        float m = INFINITY;
        for (int p = 0; p < max_paths; ++p)
            if (valid_b[p]) m = fminf(m, tau_b[p]);
        tau0 = m < INFINITY ? (double)m : 0.0;
    }
    float acc_re[kPerLane], acc_im[kPerLane];
#pragma unroll
    for (int i = 0; i < kPerLane; ++i) {
        acc_re[i] = 0.0f;
        acc_im[i] = 0.0f;
    }
    for (int p = 0; p < max_paths; ++p) {
        if (!valid_b[p]) continue;
        // tau - tau0 in float32, as the reference normalizes its float32 delays
        const double t = (double)((float)((double)tau_b[p] - tau0));
        const size_t ai = ((size_t)b * max_paths + p) * time_steps + s;
        const float ar = a[2 * ai], aim = a[2 * ai + 1];
        if (frequencies) {
#pragma unroll
            for (int i = 0; i < kPerLane; ++i) {
                const int k = k0 + lane + 64 * i;
                if (k < freq_count) {
                    float c, sn;
                    phasor(frequencies[k] * t, c, sn);
                    acc_re[i] = fmaf(ar, c, fmaf(-aim, sn, acc_re[i]));
                    acc_im[i] = fmaf(ar, sn, fmaf(aim, c, acc_im[i]));
                }
            }
        } else {
            float c = 1.0f, sn = 0.0f, rc, rs;
            phasor(64.0 * df * t, rc, rs);
#pragma unroll
            for (int i = 0; i < kPerLane; ++i) {
                if (i % kReseed == 0) phasor((f0 + df * (double)(k0 + lane + 64 * i)) * t, c, sn);
                acc_re[i] = fmaf(ar, c, fmaf(-aim, sn, acc_re[i]));
                acc_im[i] = fmaf(ar, sn, fmaf(aim, c, acc_im[i]));
                const float nc = c * rc - sn * rs;
                const float ns = c * rs + sn * rc;
                c = nc;
                sn = ns;
            }
        }
    }
    float* o = out + 2 * ((size_t)row * freq_count);
#pragma unroll
    for (int i = 0; i < kPerLane; ++i) {
        const int k = k0 + lane + 64 * i;
        if (k < freq_count) {
            o[2 * k] = acc_re[i];
            o[2 * k + 1] = acc_im[i];
        }
    }
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_paths_cfr(const float* a, const float* tau, const unsigned char* valid, int pairs, int max_paths,
                            int time_steps, const double* frequencies, double f0, double df, int freq_count,
                            int normalize_delays, float* out, void* stream) {
    if (pairs < 0 || max_paths < 0 || time_steps <= 0 || freq_count < 0)
        return wc_fail("wc_paths_cfr: invalid sizes (pairs %d, paths %d, time steps %d, frequencies %d)", pairs,
                       max_paths, time_steps, freq_count);
    if (pairs == 0 || freq_count == 0) return 0;
    if (!out) return wc_fail("wc_paths_cfr: out is null");
    if (max_paths > 0 && (!a || !tau || !valid)) return wc_fail("wc_paths_cfr: path tensors are null");
    const int blocks_per_row = (freq_count + kBlock - 1) / kBlock;
    const long long waves = (long long)pairs * time_steps * blocks_per_row;
    const long long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffLL) return wc_fail("wc_paths_cfr: launch too large (%lld waves)", waves);
    hipLaunchKernelGGL(cfr_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, a, tau, valid, pairs,
                       max_paths, time_steps, frequencies, f0, df, freq_count, blocks_per_row, normalize_delays,
                       out);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_paths_cfr", err);
}
