// Chain planning on the device: the distinct group-sequence prefixes of every
// transmitter's discovery rays, depth by depth, in (transmitter,
// lexicographic) order, and the chain table built from them.
//
// Depth d keys each ray whose first d bounces hit groups by (the id of its
// (d-1)-prefix, its d-th group): with prefix ids numbered in (transmitter,
// lexicographic) order, sorting the keys puts the d-prefixes in the same order
// (for d = 1 the parent id is the transmitter). Equal keys are one prefix; a
// scan of the run heads numbers them, and every ray keeps its prefix id for
// the next depth. A prefix is stored as (parent id, last group) per depth and
// expanded into its full sequence when the table is written.
#include "wc_api_internal.h"
#include "../../include/wcmi.h"

namespace wc {
namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int plan_grid(long long n) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > (1 << 20)) b = 1 << 20;
    return (int)b;
}

constexpr int kPlanDepths = WC_PLAN_DEPTHS;
constexpr unsigned long long kNoKey = ~0ull;

// key of ray i at depth d (kNoKey: no d-prefix)
__device__ __forceinline__ unsigned long long prefix_key(const int* __restrict__ rays, const int* __restrict__ pid,
                                                         long long i, int ray_count, int max_depth, int d,
                                                         int group_count) {
    const int parent = d == 1 ? (int)(i / ray_count) : pid[i];
    const int g = rays[i * max_depth + (d - 1)];
    if (parent < 0 || g < 0 || g >= group_count) return kNoKey;
    return (unsigned long long)parent * (unsigned long long)group_count + (unsigned long long)g;
}

__global__ void plan_keys_kernel(const int* __restrict__ rays, const int* __restrict__ pid, long long n,
                                 int ray_count, int max_depth, int d, int group_count, unsigned* __restrict__ lo,
                                 int* __restrict__ vals) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)blockDim.x * gridDim.x) {
        lo[i] = (unsigned)(prefix_key(rays, pid, i, ray_count, max_depth, d, group_count) & 0xffffffffull);
        vals[i] = (int)i;
    }
}

__global__ void plan_high_kernel(const int* __restrict__ rays, const int* __restrict__ pid, long long n,
                                 int ray_count, int max_depth, int d, int group_count, const int* __restrict__ vals,
                                 unsigned* __restrict__ hi) {
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n; s += (long long)blockDim.x * gridDim.x)
        hi[s] = (unsigned)(prefix_key(rays, pid, vals[s], ray_count, max_depth, d, group_count) >> 32);
}

__global__ void plan_heads_kernel(const int* __restrict__ rays, const int* __restrict__ pid, long long n,
                                  int ray_count, int max_depth, int d, int group_count, const int* __restrict__ vals,
                                  int* __restrict__ head) {
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n; s += (long long)blockDim.x * gridDim.x) {
        const unsigned long long k = prefix_key(rays, pid, vals[s], ray_count, max_depth, d, group_count);
        bool h = k != kNoKey;
        if (h && s > 0) h = prefix_key(rays, pid, vals[s - 1], ray_count, max_depth, d, group_count) != k;
        head[s] = h ? 1 : 0;
    }
}

// prefix id of every ray for the next depth; per prefix its parent id, last
// group and transmitter; prefixes per (transmitter, depth)
__global__ void plan_assign_kernel(const int* __restrict__ rays, const int* __restrict__ pid, long long n,
                                   int ray_count, int max_depth, int d, int group_count, const int* __restrict__ vals,
                                   const int* __restrict__ head, const int* __restrict__ first,
                                   const int* __restrict__ parent_tx, int* __restrict__ pid_next,
                                   int* __restrict__ prefix_parent, int* __restrict__ prefix_group,
                                   int* __restrict__ prefix_tx, int* __restrict__ counts) {
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < n; s += (long long)blockDim.x * gridDim.x) {
        const int i = vals[s];
        const unsigned long long k = prefix_key(rays, pid, i, ray_count, max_depth, d, group_count);
        if (k == kNoKey) {
            pid_next[i] = -1;
            continue;
        }
        const int id = first[s] + head[s] - 1;
        pid_next[i] = id;
        if (head[s]) {
            const int parent = (int)(k / (unsigned long long)group_count);
            const int tx = d == 1 ? parent : parent_tx[parent];
            prefix_parent[id] = parent;
            prefix_group[id] = (int)(k % (unsigned long long)group_count);
            prefix_tx[id] = tx;
            atomicAdd(&counts[tx * kPlanDepths + (d - 1)], 1);
        }
    }
}

// prefix id -> chain row: groups (expanded through the parents), transmitter
// and depth; row = row_start[tx] + id
__global__ void plan_table_kernel(wc_plan_levels levels, int d, int prefix_count, const int* __restrict__ row_start,
                                  int* __restrict__ groups, int* __restrict__ chain_tx, int* __restrict__ chain_depth) {
    for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < prefix_count; id += blockDim.x * gridDim.x) {
        const int tx = levels.prefix_tx[d - 1][id];
        const long long row = (long long)row_start[tx] + id;
        int seq[kPlanDepths];
#pragma unroll
        for (int k = 0; k < kPlanDepths; ++k) seq[k] = -1;
        int p = id;
        for (int k = d; k >= 1; --k) {
            seq[k - 1] = levels.prefix_group[k - 1][p];
            p = levels.prefix_parent[k - 1][p];
        }
#pragma unroll
        for (int k = 0; k < kPlanDepths; ++k) groups[row * kPlanDepths + k] = seq[k];
        chain_tx[row] = tx;
        chain_depth[row] = d;
    }
}

// every transmitter's copy of one depth's exhaustive sequence list
__global__ void plan_exhaustive_kernel(const int* __restrict__ sequences, int count, int d, int tx_count,
                                       const int* __restrict__ row_start, int* __restrict__ groups,
                                       int* __restrict__ chain_tx, int* __restrict__ chain_depth) {
    const long long n = (long long)count * tx_count;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)blockDim.x * gridDim.x) {
        const int tx = (int)(i / count);
        const int j = (int)(i - (long long)tx * count);
        const long long row = (long long)row_start[tx] + j;
#pragma unroll
        for (int k = 0; k < kPlanDepths; ++k) groups[row * kPlanDepths + k] = k < d ? sequences[(size_t)j * d + k] : -1;
        chain_tx[row] = tx;
        chain_depth[row] = d;
    }
}

__global__ void plan_zero_kernel(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0;
}

struct PlanLayout {
    size_t lo, hi, vals, keys_tmp, vals_tmp, hist, head, first, pid0, pid1, scan, total;
};

PlanLayout plan_layout(long long n) {
    const size_t m = (size_t)std::max<long long>(n, 1);
    PlanLayout l{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += align256(bytes);
        return at;
    };
    l.lo = take(4 * m);
    l.hi = take(4 * m);
    l.vals = take(4 * m);
    l.keys_tmp = take(4 * m);
    l.vals_tmp = take(4 * m);
    l.hist = take(4 * radix_sort_hist_ints((int)m));
    l.head = take(4 * m);
    l.first = take(4 * (m + 1));
    l.pid0 = take(4 * m);
    l.pid1 = take(4 * m);
    l.scan = take(scan_exclusive_workspace_bytes((long long)m));
    l.total = o;
    return l;
}

int check_level_args(const int* rays, int tx_count, int ray_count, int max_depth, int d, int group_count,
                     const char* who) {
    if (!rays) return wc_fail("%s: rays are null", who);
    if (tx_count <= 0 || ray_count <= 0) return wc_fail("%s: need transmitters and rays", who);
    if (max_depth < 1 || max_depth > kPlanDepths || d < 1 || d > max_depth)
        return wc_fail("%s: depth %d of %d outside 1..%d", who, d, max_depth, kPlanDepths);
    if (group_count <= 0) return wc_fail("%s: group_count must be positive", who);
    if ((long long)tx_count * ray_count >= (1LL << 31)) return wc_fail("%s: more than 2^31 rays", who);
    return 0;
}

}  // namespace
}  // namespace wc

using namespace wc;

extern "C" size_t wc_plan_workspace_bytes(int tx_count, int ray_count) {
    return plan_layout((long long)std::max(tx_count, 0) * std::max(ray_count, 0)).total;
}

extern "C" int wc_plan_level(const int* rays, int tx_count, int ray_count, int max_depth, int d, int group_count,
                             int* counts, long long* prefix_count, void* workspace, void* stream) {
    if (int rc = check_level_args(rays, tx_count, ray_count, max_depth, d, group_count, "wc_plan_level")) return rc;
    if (!counts || !prefix_count || !workspace) return wc_fail("wc_plan_level: null argument");
    const long long n = (long long)tx_count * ray_count;
    if ((unsigned long long)n * (unsigned long long)group_count >= (1ull << 63))
        return wc_fail("wc_plan_level: prefix keys exceed 63 bits");
    const PlanLayout l = plan_layout(n);
    char* ws = static_cast<char*>(workspace);
    unsigned* lo = reinterpret_cast<unsigned*>(ws + l.lo);
    unsigned* hi = reinterpret_cast<unsigned*>(ws + l.hi);
    int* vals = reinterpret_cast<int*>(ws + l.vals);
    unsigned* keys_tmp = reinterpret_cast<unsigned*>(ws + l.keys_tmp);
    int* vals_tmp = reinterpret_cast<int*>(ws + l.vals_tmp);
    int* hist = reinterpret_cast<int*>(ws + l.hist);
    int* head = reinterpret_cast<int*>(ws + l.head);
    int* first = reinterpret_cast<int*>(ws + l.first);
    // the parent ids of depth d live in pid[d & 1], the next depth's in the other
    const int* pid = reinterpret_cast<const int*>(ws + ((d & 1) ? l.pid1 : l.pid0));
    hipStream_t st = (hipStream_t)stream;
    if (d == 1)
        hipLaunchKernelGGL(plan_zero_kernel, dim3(plan_grid((long long)tx_count * kPlanDepths)), dim3(256), 0, st,
                           counts, tx_count * kPlanDepths);
    hipLaunchKernelGGL(plan_keys_kernel, dim3(plan_grid(n)), dim3(256), 0, st, rays, pid, n, ray_count, max_depth, d,
                       group_count, lo, vals);
    if (radix_sort_pairs(lo, vals, keys_tmp, vals_tmp, hist, (int)n, 32, st) != 0)
        return wc_fail_hip("wc_plan_level: radix sort (low word)", hipGetLastError());
    hipLaunchKernelGGL(plan_high_kernel, dim3(plan_grid(n)), dim3(256), 0, st, rays, pid, n, ray_count, max_depth, d,
                       group_count, vals, hi);
    if (radix_sort_pairs(hi, vals, keys_tmp, vals_tmp, hist, (int)n, 32, st) != 0)
        return wc_fail_hip("wc_plan_level: radix sort (high word)", hipGetLastError());
    hipLaunchKernelGGL(plan_heads_kernel, dim3(plan_grid(n)), dim3(256), 0, st, rays, pid, n, ray_count, max_depth, d,
                       group_count, vals, head);
    if (scan_exclusive(head, (int)n, first, ws + l.scan, prefix_count, st) != 0)
        return wc_fail_hip("wc_plan_level: scan", hipGetLastError());
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_plan_level", err);
}

extern "C" int wc_plan_assign(const int* rays, int tx_count, int ray_count, int max_depth, int d, int group_count,
                              const int* parent_tx, int* prefix_parent, int* prefix_group, int* prefix_tx,
                              int* counts, void* workspace, void* stream) {
    if (int rc = check_level_args(rays, tx_count, ray_count, max_depth, d, group_count, "wc_plan_assign")) return rc;
    if (!counts || !workspace || (d > 1 && !parent_tx)) return wc_fail("wc_plan_assign: null argument");
    const long long n = (long long)tx_count * ray_count;
    const PlanLayout l = plan_layout(n);
    char* ws = static_cast<char*>(workspace);
    const int* vals = reinterpret_cast<const int*>(ws + l.vals);
    const int* head = reinterpret_cast<const int*>(ws + l.head);
    const int* first = reinterpret_cast<const int*>(ws + l.first);
    const int* pid = reinterpret_cast<const int*>(ws + ((d & 1) ? l.pid1 : l.pid0));
    int* pid_next = reinterpret_cast<int*>(ws + ((d & 1) ? l.pid0 : l.pid1));
    hipLaunchKernelGGL(plan_assign_kernel, dim3(plan_grid(n)), dim3(256), 0, (hipStream_t)stream, rays, pid, n,
                       ray_count, max_depth, d, group_count, vals, head, first, parent_tx, pid_next, prefix_parent,
                       prefix_group, prefix_tx, counts);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_plan_assign", err);
}

extern "C" int wc_plan_table(const wc_plan_levels* levels, int d, int prefix_count, const int* row_start,
                             int* groups, int* chain_tx, int* chain_depth, void* stream) {
    if (!levels || d < 1 || d > kPlanDepths || prefix_count < 0) return wc_fail("wc_plan_table: bad arguments");
    if (prefix_count == 0) return 0;
    if (!row_start || !groups || !chain_tx || !chain_depth) return wc_fail("wc_plan_table: null argument");
    for (int k = 0; k < d; ++k)
        if (!levels->prefix_parent[k] || !levels->prefix_group[k] || !levels->prefix_tx[k])
            return wc_fail("wc_plan_table: level %d of depth %d is missing", k + 1, d);
    hipLaunchKernelGGL(plan_table_kernel, dim3(plan_grid(prefix_count)), dim3(256), 0, (hipStream_t)stream, *levels, d,
                       prefix_count, row_start, groups, chain_tx, chain_depth);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_plan_table", err);
}

extern "C" int wc_plan_exhaustive(const int* sequences, int count, int d, int tx_count, const int* row_start,
                                  int* groups, int* chain_tx, int* chain_depth, void* stream) {
    if (d < 1 || d > kPlanDepths || count < 0 || tx_count < 0) return wc_fail("wc_plan_exhaustive: bad arguments");
    if (count == 0 || tx_count == 0) return 0;
    if (!sequences || !row_start || !groups || !chain_tx || !chain_depth)
        return wc_fail("wc_plan_exhaustive: null argument");
    hipLaunchKernelGGL(plan_exhaustive_kernel, dim3(plan_grid((long long)count * tx_count)), dim3(256), 0,
                       (hipStream_t)stream, sequences, count, d, tx_count, row_start, groups, chain_tx, chain_depth);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_plan_exhaustive", err);
}
