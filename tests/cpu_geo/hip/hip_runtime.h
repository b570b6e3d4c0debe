// Host shim for CPU checks: lets the device sources compile as plain C++ so
// their geometric tests (culls, frames, triangle tests) run on the host.
// Kernels become ordinary functions that nothing calls, launches are no-ops,
// and the few wavefront intrinsics the sources name act on a single lane.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __constant__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(x)

typedef int hipError_t;
typedef void* hipStream_t;
#define hipSuccess 0

struct float4 { float x, y, z, w; };
struct int4 { int x, y, z, w; };
struct uint4 { unsigned x, y, z, w; };
struct float2 { float x, y; };
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline int4 make_int4(int x, int y, int z, int w) { return {x, y, z, w}; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline void wc_host_sincosf(float x, float* s, float* c) { *s = std::sin(x); *c = std::cos(x); }
#define __sincosf wc_host_sincosf
using std::max;
using std::min;

struct wc_dim3 { unsigned x = 0, y = 0, z = 0; };
inline wc_dim3 threadIdx, blockIdx, blockDim, gridDim;
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
#define hipLaunchKernelGGL(...) ((void)0)
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
enum { hipDeviceAttributeMultiprocessorCount = 0 };
inline hipError_t hipDeviceGetAttribute(int* v, int, int) { *v = 256; return 0; }
inline hipError_t hipMemsetAsync(void*, int, size_t, hipStream_t) { return 0; }

template <class T> inline T __shfl(T v, int) { return v; }
template <class T> inline T __shfl_up(T v, int) { return v; }
template <class T> inline T __shfl_xor(T v, int) { return v; }
inline void __syncthreads() {}
inline unsigned long long __ballot(int v) { return v ? 1ull : 0ull; }
inline int __ffs(unsigned x) { return x ? __builtin_ctz(x) + 1 : 0; }
inline int __ffs(int x) { return __ffs((unsigned)x); }
inline int __ffsll(unsigned long long x) { return x ? __builtin_ctzll(x) + 1 : 0; }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned atomicOr(unsigned* p, unsigned v) { unsigned o = *p; *p |= v; return o; }
inline int atomicOr(int* p, int v) { int o = *p; *p |= v; return o; }
inline int atomicAdd(int* p, int v) { int o = *p; *p += v; return o; }
inline unsigned atomicAdd(unsigned* p, unsigned v) { unsigned o = *p; *p += v; return o; }
inline double unsafeAtomicAdd(double* p, double v) { double o = *p; *p += v; return o; }
#define __builtin_amdgcn_wave_barrier() ((void)0)
#define __builtin_amdgcn_fence(order, scope) ((void)0)
