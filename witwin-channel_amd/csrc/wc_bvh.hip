// LBVH construction on the GPU: 30-bit Morton codes of triangle centroids,
// a stable LSD radix sort with LDS-ranked scatter (8-bit digits, 4 passes),
// Karras' radix-tree hierarchy (HPG 2012) and a bottom-up box pass with
// arrival counters. Output nodes keep both child boxes inline so one 64 B
// load tests two children during traversal.
#include "wc_api_internal.h"
#include "wc_scene.h"

#include <cfloat>

namespace wc {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

__global__ void centroid_bounds_kernel(const float* __restrict__ vertices, const int* __restrict__ faces, int n,
                                       float* __restrict__ partial) {
    __shared__ float smin[3][256];
    __shared__ float smax[3][256];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const int a = faces[3 * i], b = faces[3 * i + 1], c = faces[3 * i + 2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float cen = (vertices[3 * a + k] + vertices[3 * b + k] + vertices[3 * c + k]) * (1.0f / 3.0f);
            lo[k] = fminf(lo[k], cen);
            hi[k] = fmaxf(hi[k], cen);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        smin[k][threadIdx.x] = lo[k];
        smax[k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int stride = blockDim.x / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                smin[k][threadIdx.x] = fminf(smin[k][threadIdx.x], smin[k][threadIdx.x + stride]);
                smax[k][threadIdx.x] = fmaxf(smax[k][threadIdx.x], smax[k][threadIdx.x + stride]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            partial[blockIdx.x * 6 + k] = smin[k][0];
            partial[blockIdx.x * 6 + 3 + k] = smax[k][0];
        }
    }
}

__device__ __forceinline__ unsigned int expand_bits(unsigned int v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void morton_kernel(const float* __restrict__ vertices, const int* __restrict__ faces, int n,
                              const float* __restrict__ partial, int partial_count, unsigned int* __restrict__ keys,
                              int* __restrict__ values) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int p = 0; p < partial_count; ++p) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], partial[p * 6 + k]);
            hi[k] = fmaxf(hi[k], partial[p * 6 + 3 + k]);
        }
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) {
        const int a = faces[3 * i], b = faces[3 * i + 1], c = faces[3 * i + 2];
        unsigned int code = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float cen = (vertices[3 * a + k] + vertices[3 * b + k] + vertices[3 * c + k]) * (1.0f / 3.0f);
            const float ext = hi[k] - lo[k];
            float u = ext > 0.0f ? (cen - lo[k]) / ext : 0.5f;
            u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
            code |= expand_bits((unsigned int)u) << (2 - k);
        }
        keys[i] = code;
        values[i] = i;
    }
}

// ---- stable LSD radix sort (uint32 keys, int values) -------------------------

__global__ void radix_histogram_kernel(const unsigned int* __restrict__ keys, int n, int shift,
                                       int* __restrict__ hist, int blocks) {
    __shared__ int counts[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += blockDim.x) counts[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int idx = base + it * kSortThreads + threadIdx.x;
        if (idx < n) atomicAdd(&counts[(keys[idx] >> shift) & (kRadix - 1)], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += blockDim.x) hist[d * blocks + blockIdx.x] = counts[d];
}

// exclusive scan of hist (digit-major, block-minor) in one workgroup
__global__ void radix_scan_kernel(int* __restrict__ hist, int count) {
    __shared__ int partial[1024];
    const int tid = threadIdx.x;
    const int per = (count + blockDim.x - 1) / blockDim.x;
    const int begin = tid * per;
    const int end = min(begin + per, count);
    int sum = 0;
    for (int i = begin; i < end; ++i) sum += hist[i];
    partial[tid] = sum;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            const int v = partial[i];
            partial[i] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = partial[tid];
    for (int i = begin; i < end; ++i) {
        const int v = hist[i];
        hist[i] = run;
        run += v;
    }
}

// Stable scatter: tile elements are ranked in index order (item-major,
// thread-minor); within a wavefront the rank among equal digits comes from
// eight ballots, across wavefronts from per-wave digit counts in LDS.
__global__ void radix_scatter_kernel(const unsigned int* __restrict__ keys_in, const int* __restrict__ values_in,
                                     unsigned int* __restrict__ keys_out, int* __restrict__ values_out, int n,
                                     int shift, const int* __restrict__ hist, int blocks) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ int offsets[kRadix];
    __shared__ int wave_counts[kWaves][kRadix];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    for (int d = tid; d < kRadix; d += blockDim.x) offsets[d] = hist[d * blocks + blockIdx.x];
    const int base = blockIdx.x * kSortTile;
    const unsigned long long lane_lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int it = 0; it < kSortItems; ++it) {
        for (int d = tid; d < kWaves * kRadix; d += blockDim.x) (&wave_counts[0][0])[d] = 0;
        __syncthreads();
        const int idx = base + it * kSortThreads + tid;
        const bool active = idx < n;
        unsigned int key = active ? keys_in[idx] : 0u;
        const int digit = (int)((key >> shift) & (kRadix - 1));
        unsigned long long match = __ballot(active);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const unsigned long long bal = __ballot(active && ((digit >> b) & 1));
            match &= ((digit >> b) & 1) ? bal : ~bal;
        }
        const int lane_rank = __popcll(match & lane_lt);
        // the lowest matching lane publishes the wave's count for this digit
        if (active && lane_rank == 0) wave_counts[wave][digit] = __popcll(match);
        __syncthreads();
        if (active) {
            int before = 0;
            for (int w = 0; w < wave; ++w) before += wave_counts[w][digit];
            const int dst = offsets[digit] + before + lane_rank;
            keys_out[dst] = key;
            values_out[dst] = values_in[idx];
        }
        __syncthreads();
        // advance the per-digit offsets by this item's block totals
        for (int d = tid; d < kRadix; d += blockDim.x) {
            int total = 0;
            for (int w = 0; w < kWaves; ++w) total += wave_counts[w][d];
            offsets[d] += total;
        }
        __syncthreads();
    }
}

// ---- Karras hierarchy ----------------------------------------------------------

__device__ __forceinline__ int delta(const unsigned int* keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned int ki = keys[i], kj = keys[j];
    if (ki == kj) return 32 + __clz((unsigned int)(i ^ j));
    return __clz(ki ^ kj);
}

__global__ void karras_kernel(const unsigned int* __restrict__ keys, int n, int* __restrict__ child,
                              int* __restrict__ parent) {
    // child[2*i], child[2*i+1] : >= 0 internal, < 0 leaf ~k ; parent of internal i at parent[i],
    // parent of leaf k at parent[n - 1 + k]
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += blockDim.x * gridDim.x) {
        const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(keys, n, i, i - d);
        int lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1) {
            if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        }
        const int j = i + l * d;
        const int dnode = delta(keys, n, i, j);
        int s = 0;
        int t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = i + s * d + min(d, 0);
        const int left = min(i, j) == gamma ? ~gamma : gamma;
        const int right = max(i, j) == gamma + 1 ? ~(gamma + 1) : gamma + 1;
        child[2 * i] = left;
        child[2 * i + 1] = right;
        if (left < 0) parent[n - 1 + ~left] = i; else parent[left] = i;
        if (right < 0) parent[n - 1 + ~right] = i; else parent[right] = i;
    }
}

__global__ void gather_tris_kernel(const float* __restrict__ vertices, const int* __restrict__ faces,
                                   const int* __restrict__ face_group, const int* __restrict__ face_material,
                                   const int* __restrict__ order, int n, Tri* __restrict__ sorted_tris,
                                   Tri* __restrict__ face_tris, float* __restrict__ leaf_box) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x) {
        const int f = order[k];
        const int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
        Tri t;
        t.v0 = make_float4(vertices[3 * a], vertices[3 * a + 1], vertices[3 * a + 2], __int_as_float(face_group[f]));
        t.v1 = make_float4(vertices[3 * b], vertices[3 * b + 1], vertices[3 * b + 2], __int_as_float(f));
        t.v2 = make_float4(vertices[3 * c], vertices[3 * c + 1], vertices[3 * c + 2], __int_as_float(face_material[f]));
        sorted_tris[k] = t;
        face_tris[f] = t;
        float* box = leaf_box + 6 * k;
        box[0] = fminf(fminf(t.v0.x, t.v1.x), t.v2.x);
        box[1] = fminf(fminf(t.v0.y, t.v1.y), t.v2.y);
        box[2] = fminf(fminf(t.v0.z, t.v1.z), t.v2.z);
        box[3] = fmaxf(fmaxf(t.v0.x, t.v1.x), t.v2.x);
        box[4] = fmaxf(fmaxf(t.v0.y, t.v1.y), t.v2.y);
        box[5] = fmaxf(fmaxf(t.v0.z, t.v1.z), t.v2.z);
    }
}

// Bottom-up boxes: each leaf walks to the root; the second thread to reach
// a node (arrival counter) merges both children and continues.
__global__ void refit_kernel(int n, const int* __restrict__ child, const int* __restrict__ parent,
                             const float* __restrict__ leaf_box, float* __restrict__ node_box,
                             int* __restrict__ arrival, Node* __restrict__ nodes) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += blockDim.x * gridDim.x) {
        int node = parent[n - 1 + k];
        while (node >= 0) {
            __threadfence();
            const int prev = atomicAdd(&arrival[node], 1);
            if (prev == 0) break;
            __threadfence();
            const int c0 = child[2 * node], c1 = child[2 * node + 1];
            float b0[6], b1[6];
            const float* p0 = c0 < 0 ? leaf_box + 6 * (~c0) : node_box + 6 * c0;
            const float* p1 = c1 < 0 ? leaf_box + 6 * (~c1) : node_box + 6 * c1;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                // device-coherent loads: the sibling box was written by another workgroup
                b0[q] = __hip_atomic_load(p0 + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                b1[q] = __hip_atomic_load(p1 + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            float* out = node_box + 6 * node;
            out[0] = fminf(b0[0], b1[0]); out[1] = fminf(b0[1], b1[1]); out[2] = fminf(b0[2], b1[2]);
            out[3] = fmaxf(b0[3], b1[3]); out[4] = fmaxf(b0[4], b1[4]); out[5] = fmaxf(b0[5], b1[5]);
            Node nd;
            nd.a = make_float4(b0[0], b0[1], b0[2], b0[3]);
            nd.b = make_float4(b0[4], b0[5], b1[0], b1[1]);
            nd.c = make_float4(b1[2], b1[3], b1[4], b1[5]);
            nd.d = make_int4(c0, c1, 0, 0);
            nodes[node] = nd;
            node = parent[node];
        }
    }
}

__global__ void single_leaf_kernel(const float* __restrict__ leaf_box, Node* __restrict__ nodes) {
    Node nd;
    nd.a = make_float4(leaf_box[0], leaf_box[1], leaf_box[2], leaf_box[3]);
    // second child: inverted (empty) box never hit
    nd.b = make_float4(leaf_box[4], leaf_box[5], FLT_MAX, FLT_MAX);
    nd.c = make_float4(FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX);
    nd.d = make_int4(~0, ~0, 0, 0);
    nodes[0] = nd;
}

__global__ void fill_int_kernel(int* p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) p[i] = v;
}

int grid_for(int n, int threads) { return n <= 0 ? 1 : std::min((n + threads - 1) / threads, 65535 * 4); }

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

struct BuildWorkspace {
    unsigned int* keys0;
    unsigned int* keys1;
    int* vals0;
    int* vals1;
    int* hist;
    float* partial;
    int* child;
    int* parent;
    int* arrival;
    float* leaf_box;
    float* node_box;
};

size_t layout(int n, BuildWorkspace* w, char* base) {
    const int blocks = std::max(1, (n + kSortTile - 1) / kSortTile);
    const int partials = 256;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align_up(bytes);
        return p;
    };
    BuildWorkspace ws;
    ws.keys0 = (unsigned int*)take(sizeof(unsigned int) * std::max(n, 1));
    ws.keys1 = (unsigned int*)take(sizeof(unsigned int) * std::max(n, 1));
    ws.vals0 = (int*)take(sizeof(int) * std::max(n, 1));
    ws.vals1 = (int*)take(sizeof(int) * std::max(n, 1));
    ws.hist = (int*)take(sizeof(int) * kRadix * blocks);
    ws.partial = (float*)take(sizeof(float) * 6 * partials);
    ws.child = (int*)take(sizeof(int) * 2 * std::max(n, 1));
    ws.parent = (int*)take(sizeof(int) * 2 * std::max(n, 1));
    ws.arrival = (int*)take(sizeof(int) * std::max(n, 1));
    ws.leaf_box = (float*)take(sizeof(float) * 6 * std::max(n, 1));
    ws.node_box = (float*)take(sizeof(float) * 6 * std::max(n, 1));
    if (w) *w = ws;
    return off;
}

}  // namespace

// Radix-sort helper also used by the export path (stable, uint32 keys).
int radix_sort_pairs(unsigned int* keys, int* values, unsigned int* keys_tmp, int* values_tmp, int* hist, int n,
                     int key_bits, hipStream_t stream) {
    if (n <= 1) return 0;
    const int blocks = (n + kSortTile - 1) / kSortTile;
    unsigned int* kin = keys;
    unsigned int* kout = keys_tmp;
    int* vin = values;
    int* vout = values_tmp;
    int passes = (key_bits + kRadixBits - 1) / kRadixBits;
    if (passes % 2 == 1) passes += 1;  // even pass count leaves the result in `keys`
    for (int p = 0; p < passes; ++p) {
        const int shift = p * kRadixBits;
        hipLaunchKernelGGL(radix_histogram_kernel, dim3(blocks), dim3(kSortThreads), 0, stream, kin, n, shift, hist,
                           blocks);
        hipLaunchKernelGGL(radix_scan_kernel, dim3(1), dim3(1024), 0, stream, hist, kRadix * blocks);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(blocks), dim3(kSortThreads), 0, stream, kin, vin, kout, vout, n,
                           shift, hist, blocks);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t radix_sort_hist_ints(int n) { return (size_t)kRadix * std::max(1, (n + kSortTile - 1) / kSortTile); }

}  // namespace wc

using namespace wc;

extern "C" size_t wc_bvh_workspace_bytes(int face_count) { return layout(face_count, nullptr, nullptr); }

extern "C" int wc_bvh_build(const float* vertices, const int* faces, const int* face_group, const int* face_material,
                            int face_count, void* nodes_out, void* sorted_tris_out, void* face_tris_out,
                            void* workspace, size_t workspace_bytes, hipStream_t stream) {
    if (face_count < 0) return wc_fail("wc_bvh_build: face_count must be non-negative");
    if (face_count == 0) return 0;
    if (!vertices || !faces || !face_group || !face_material || !nodes_out || !sorted_tris_out || !face_tris_out ||
        !workspace)
        return wc_fail("wc_bvh_build: null pointer argument");
    if (workspace_bytes < wc_bvh_workspace_bytes(face_count))
        return wc_fail("wc_bvh_build: workspace is smaller than wc_bvh_workspace_bytes(face_count)");
    BuildWorkspace ws;
    layout(face_count, &ws, (char*)workspace);
    const int n = face_count;
    Node* nodes = (Node*)nodes_out;
    const int partials = std::min(256, grid_for(n, 256));
    hipLaunchKernelGGL(centroid_bounds_kernel, dim3(partials), dim3(256), 0, stream, vertices, faces, n, ws.partial);
    hipLaunchKernelGGL(morton_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, vertices, faces, n, ws.partial,
                       partials, ws.keys0, ws.vals0);
    if (radix_sort_pairs(ws.keys0, ws.vals0, ws.keys1, ws.vals1, ws.hist, n, 32, stream) != 0)
        return wc_fail("wc_bvh_build: radix sort launch failed");
    hipLaunchKernelGGL(gather_tris_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, vertices, faces, face_group,
                       face_material, ws.vals0, n, (Tri*)sorted_tris_out, (Tri*)face_tris_out, ws.leaf_box);
    if (n == 1) {
        hipLaunchKernelGGL(single_leaf_kernel, dim3(1), dim3(1), 0, stream, ws.leaf_box, nodes);
    } else {
        hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(2 * n, 256)), dim3(256), 0, stream, ws.parent, 2 * n, -1);
        hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, ws.arrival, n, 0);
        hipLaunchKernelGGL(karras_kernel, dim3(grid_for(n - 1, 256)), dim3(256), 0, stream, ws.keys0, n, ws.child,
                           ws.parent);
        hipLaunchKernelGGL(refit_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, n, ws.child, ws.parent,
                           ws.leaf_box, ws.node_box, ws.arrival, nodes);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return wc_fail_hip("wc_bvh_build", err);
    return 0;
}

extern "C" int wc_bvh_node_count(int face_count) { return face_count <= 1 ? 1 : face_count - 1; }
