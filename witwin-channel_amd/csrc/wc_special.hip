// Element-wise evaluation of the device special functions behind the UTD
// coefficient (transition function, complex Fresnel integral), so that tests
// can check the exact code the radiomap kernels run.
#include "wc_api_internal.h"
#include "wc_utd.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

__global__ void special_kernel(int kind, const double* __restrict__ x, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double re = 0.0, im = 0.0;
    if (kind == 0) {
        utd_transition(x[i], re, im);
    } else {
        fresnel_complex(x[i], re, im);
    }
    out[2 * i] = re;
    out[2 * i + 1] = im;
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" int wc_utd_special_eval(int kind, const double* x, int n, double* out, void* stream) {
    if (kind != 0 && kind != 1) return wc_fail("wc_utd_special_eval: kind must be 0 (transition) or 1 (Fresnel)");
    if (n < 0) return wc_fail("wc_utd_special_eval: n must be non-negative");
    if (n == 0) return 0;
    if (!x || !out) return wc_fail("wc_utd_special_eval: null argument");
    hipLaunchKernelGGL(special_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, kind, x, n, out);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_utd_special_eval", err);
}
