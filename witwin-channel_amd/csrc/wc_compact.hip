// Canonical path compaction on the device: the canonical (receiver-major,
// transmitter-minor) order of the (transmitter, receiver) cells, the row
// offsets of every pair from a scan of the per-cell path counts, the canonical
// order inside each pair (a rank sort of compact row keys staged in LDS, one
// wavefront per pair), the identity de-duplication, the per-pair path cap and
// the gather list of the kept rows.
#include "wc_api_internal.h"
#include "../../include/wcmi.h"

#include <climits>

namespace wc {
namespace {

size_t align_up256(size_t n) { return (n + 255) & ~(size_t)255; }

int grid_for(long long n, int per_block) {
    long long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > (1 << 20)) b = 1 << 20;
    return (int)b;
}

__device__ __forceinline__ int wave_scan_incl(int v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int n = __shfl_up(v, off);
        if (lane >= off) v += n;
    }
    return v;
}

// ------------------------------------------------------------------ scan
// Exclusive scan of n ints (optionally gathered through `index`) into
// out[0..n]; optionally scattered as scatter_out[index[i]] = out[i]. Tiles of
// 1,024 values (256 threads x 4); tile totals are scanned in int64 by one block,
// whose grand total is also written to *total (the caller checks it fits int).
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void __launch_bounds__(kScanThreads) scan_tile_kernel(const int* __restrict__ values,
                                                                 const int* __restrict__ index, int n,
                                                                 int* __restrict__ out,
                                                                 long long* __restrict__ tile_sum) {
    __shared__ int wave_total[kScanThreads / 64];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long first = (long long)tile * kScanTile + (long long)threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = first + k;
        v[k] = i < n ? values[index ? index[i] : i] : 0;
        sum += v[k];
    }
    const int incl = wave_scan_incl(sum, lane);
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wave_total[w];
    int run = before + incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = first + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanThreads - 1) tile_sum[tile] = (long long)run;
}

__global__ void __launch_bounds__(1024) scan_tile_sums_kernel(long long* __restrict__ tile_sum, int tiles,
                                                              long long* __restrict__ total) {
    __shared__ long long part[1024];
    long long carry = 0;
    for (int base = 0; base < tiles; base += 1024) {
        const int i = base + threadIdx.x;
        const long long x = i < tiles ? tile_sum[i] : 0;
        part[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const long long y = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < tiles) tile_sum[i] = carry + part[threadIdx.x] - x;
        const long long chunk = part[1023];
        __syncthreads();
        carry += chunk;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kScanThreads) scan_add_kernel(int n, const int* __restrict__ index,
                                                                const long long* __restrict__ tile_sum,
                                                                const long long* __restrict__ total,
                                                                int* __restrict__ out, int* __restrict__ scatter_out) {
    const long long i = (long long)blockIdx.x * kScanThreads + threadIdx.x;
    if (i < n) {
        const int v = (int)((long long)out[i] + tile_sum[i / kScanTile]);
        out[i] = v;
        if (scatter_out) scatter_out[index[i]] = v;
    }
    if (i == 0) out[n] = (int)*total;
}

size_t scan_bytes(long long n) { return align_up256(8 * (size_t)((n + kScanTile - 1) / kScanTile + 1)) + 256; }

// tile sums then the grand total at the end of the workspace slice
int scan_launch(const int* values, const int* index, int n, int* out, int* scatter_out, void* ws, long long* total,
                hipStream_t st) {
    const int tiles = std::max(1, (n + kScanTile - 1) / kScanTile);
    long long* tile_sum = static_cast<long long*>(ws);
    long long* grand = total ? total : tile_sum + tiles;
    hipLaunchKernelGGL(scan_tile_kernel, dim3(tiles), dim3(kScanThreads), 0, st, values, index, n, out, tile_sum);
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(1), dim3(1024), 0, st, tile_sum, tiles, grand);
    hipLaunchKernelGGL(scan_add_kernel, dim3(grid_for((long long)n + 1, kScanThreads)), dim3(kScanThreads), 0, st, n,
                       index, tile_sum, grand, out, scatter_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------------ pair order
// key of cell c = rx * T + tx (rx = position of the receiver in its
// transmitter's set), payload c
__global__ void pair_keys_kernel(const int* __restrict__ cell_offset, int tx_count, int cells,
                                 unsigned* __restrict__ keys, int* __restrict__ vals) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += blockDim.x * gridDim.x) {
        int lo = 0, hi = tx_count - 1;  // last t with cell_offset[t] <= c
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cell_offset[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        keys[c] = (unsigned)(c - cell_offset[lo]) * (unsigned)tx_count + (unsigned)lo;
        vals[c] = c;
    }
}

// ------------------------------------------------------------------ in-pair order
// Canonical key of a row: (depth, component), primitive sequence columns
// 0..4 (INT_MIN past the row width, equal for every row), primitive, edge
// (reference native/channel/kernels/evaluated_paths.cu:654-700, the sort key of
// the canonical compact).
constexpr int kKeyInts = 8;
constexpr int kKeySeq = 5;
constexpr int kStageRows = 256;

__device__ __forceinline__ int key_part(const wc_path_rows& r, int row, int q) {
    if (q == 0) return r.depth[row] * 256 + r.component_id[row];
    if (q <= kKeySeq) return q - 1 < r.width ? r.primitive_sequence[(size_t)row * r.width + (q - 1)] : INT_MIN;
    return q == 6 ? r.primitive_id[row] : r.edge_id[row];
}

// Rank sort, one wavefront per pair: the rank of row i is the number of rows
// before it in (key, original position) order, counted against the pair's
// keys staged in LDS 256 at a time; row i goes to position begin + rank.
// Stable, and the same permutation as any stable sort on the key.
__global__ void __launch_bounds__(64) pair_rank_kernel(wc_path_rows rows, const int* __restrict__ pair_offset,
                                                       int pair_count, int* __restrict__ perm) {
    __shared__ int stage[kKeyInts][kStageRows];
    const int lane = threadIdx.x;
    for (int p = blockIdx.x; p < pair_count; p += gridDim.x) {
        const int begin = pair_offset[p], n = pair_offset[p + 1] - begin;
        if (n <= 1) {
            if (n == 1 && lane == 0) perm[begin] = begin;
            continue;
        }
        for (int i0 = 0; i0 < n; i0 += 64) {
            const int i = i0 + lane;
            int mine[kKeyInts];
#pragma unroll
            for (int q = 0; q < kKeyInts; ++q) mine[q] = i < n ? key_part(rows, begin + i, q) : 0;
            int rank = 0;
            for (int j0 = 0; j0 < n; j0 += kStageRows) {
                const int m = min(kStageRows, n - j0);
                __syncthreads();
                for (int jj = lane; jj < m; jj += 64)
#pragma unroll
                    for (int q = 0; q < kKeyInts; ++q) stage[q][jj] = key_part(rows, begin + j0 + jj, q);
                __syncthreads();
                if (i < n) {
                    for (int jj = 0; jj < m; ++jj) {
                        int c = 0;
#pragma unroll
                        for (int q = 0; q < kKeyInts; ++q) {
                            const int a = stage[q][jj];
                            if (c == 0 && a != mine[q]) c = a < mine[q] ? -1 : 1;
                        }
                        rank += (c < 0 || (c == 0 && j0 + jj < i)) ? 1 : 0;
                    }
                }
            }
            if (i < n) perm[begin + rank] = begin + i;
        }
    }
}

// Identity of a row: per slot, its event type and object id (reference
// evaluated_paths.cu:126-310): reflection and transmission slots carry the
// primitive, the diffraction slot the edge; the two leading slots of coupled
// rows (3: reflection then diffraction, 4: diffraction then reflection, 7:
// two diffractions) carry their sequence entries.
__device__ __forceinline__ int event_type(int component, int depth, int slot, int width) {
    if (component == 1 && slot < depth) return 1;
    if (component == 2 && depth > 0 && slot == 0) return 2;
    if (component == 5 && slot < depth) return 4;
    if (component == 6 && depth > 0 && slot == 0) return 8;
    if (width >= 2 && depth >= 2 && slot < 2) {
        if (component == 3) return slot == 0 ? 1 : 2;
        if (component == 4) return slot == 0 ? 2 : 1;
        if (component == 7) return 2;
    }
    return 0;
}

__device__ __forceinline__ bool same_identity(const wc_path_rows& r, int a, int b) {
    for (int j = 0; j < r.width; ++j) {
        const int ta = event_type(r.component_id[a], r.depth[a], j, r.width);
        const int tb = event_type(r.component_id[b], r.depth[b], j, r.width);
        if (ta != tb) return false;
        if (ta == 0) continue;
        const int oa = (r.component_id[a] == 2 && j == 0) ? r.edge_id[a] : r.primitive_sequence[(size_t)a * r.width + j];
        const int ob = (r.component_id[b] == 2 && j == 0) ? r.edge_id[b] : r.primitive_sequence[(size_t)b * r.width + j];
        if (oa != ob) return false;
    }
    return true;
}

// Identity groups are runs of the sorted pair: a row starts a group unless it
// has the identity of the row before it (reference evaluated_paths.cu:216-320).
// One wavefront per pair: each group's first lane keeps the group's first
// strictly-shortest row (none when a length is NaN); the pair keeps
// min(winners, cap) rows (cap < 0: all), written to kept_count.
__global__ void __launch_bounds__(64) pair_keep_kernel(wc_path_rows rows, const int* __restrict__ pair_offset,
                                                       int pair_count, const int* __restrict__ perm,
                                                       unsigned char* __restrict__ keep, int cap,
                                                       int* __restrict__ kept_count, int* __restrict__ dropped) {
    const int lane = threadIdx.x;
    for (int p = blockIdx.x; p < pair_count; p += gridDim.x) {
        const int begin = pair_offset[p], n = pair_offset[p + 1] - begin;
        int groups = 0;
        for (int s0 = 0; s0 < n; s0 += 64) {
            const int s = s0 + lane;
            bool head = false, won = false;
            if (s < n) head = s == 0 || !same_identity(rows, perm[begin + s - 1], perm[begin + s]);
            if (head) {
                const int a = perm[begin + s];
                int e = s + 1;
                float best = rows.path_length_m[a];
                bool nan = isnan(best);
                while (e < n && same_identity(rows, a, perm[begin + e])) {
                    const float len = rows.path_length_m[perm[begin + e]];
                    nan = nan || isnan(len);
                    best = fminf(best, len);
                    ++e;
                }
                for (int x = s; x < e; ++x) {
                    const bool win = !nan && !won && rows.path_length_m[perm[begin + x]] == best;
                    keep[begin + x] = win ? 1 : 0;
                    won = won || win;
                }
            }
            groups += __popcll(__ballot(won));
        }
        if (lane == 0) {
            const int kept = cap >= 0 ? min(groups, cap) : groups;
            if (kept_count) kept_count[p] = kept;
            if (dropped && groups > kept) atomicAdd(dropped, groups - kept);
        }
    }
}

// Gather list: the kept rows of pair p, in sorted order, at kept_offset[p] ..
// kept_offset[p + 1] (source row ids); rows past the pair's cap are left out.
__global__ void __launch_bounds__(64) pair_gather_kernel(const int* __restrict__ pair_offset, int pair_count,
                                                         const int* __restrict__ perm,
                                                         const unsigned char* __restrict__ keep,
                                                         const int* __restrict__ kept_offset,
                                                         int* __restrict__ source_row) {
    const int lane = threadIdx.x;
    for (int p = blockIdx.x; p < pair_count; p += gridDim.x) {
        const int begin = pair_offset[p], n = pair_offset[p + 1] - begin;
        const int base = kept_offset[p], limit = kept_offset[p + 1] - base;
        int taken = 0;
        for (int s0 = 0; s0 < n && taken < limit; s0 += 64) {
            const int s = s0 + lane;
            const int k = (s < n && keep[begin + s]) ? 1 : 0;
            const int incl = wave_scan_incl(k, lane);
            const int rank = taken + incl - k;
            if (k && rank < limit) source_row[base + rank] = perm[begin + s];
            taken += __shfl(incl, 63);
        }
    }
}

__global__ void zero_ints_kernel(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0;
}

// workspace slices: [scan | sort keys x2 | sort values tmp | histogram | kept counts]
struct CompactLayout {
    size_t scan, keys0, keys1, vals_tmp, hist, kept, total;
};

CompactLayout compact_layout(int cells) {
    const size_t n = (size_t)std::max(cells, 1);
    CompactLayout l{};
    size_t o = 0;
    l.scan = o;
    o += scan_bytes(n + 1);
    l.keys0 = o;
    o += align_up256(4 * n);
    l.keys1 = o;
    o += align_up256(4 * n);
    l.vals_tmp = o;
    o += align_up256(4 * n);
    l.hist = o;
    o += align_up256(4 * radix_sort_hist_ints(cells));
    l.kept = o;
    o += align_up256(4 * n);
    l.total = o;
    return l;
}

int check_rows(const wc_path_rows* rows, const char* who) {
    if (!rows) return wc_fail("%s: rows are null", who);
    if (rows->width < 1 || rows->width > kKeySeq)
        return wc_fail("%s: row width %d outside 1..%d", who, rows->width, kKeySeq);
    return 0;
}

}  // namespace

size_t scan_exclusive_workspace_bytes(long long n) { return scan_bytes(n + 1); }

int scan_exclusive(const int* values, int n, int* out, void* workspace, long long* total, hipStream_t stream) {
    return scan_launch(values, nullptr, n, out, nullptr, workspace, total, stream);
}

}  // namespace wc

using namespace wc;

extern "C" size_t wc_paths_compact_workspace_bytes(int cells) { return compact_layout(cells).total; }

extern "C" int wc_paths_pair_order(const int* cell_offset, int tx_count, int cells, int max_rx, int* pair_cell,
                                   void* workspace, void* stream) {
    if (cells < 0 || tx_count < 0 || max_rx < 0) return wc_fail("wc_paths_pair_order: sizes must be non-negative");
    if (cells == 0) return 0;
    if (!cell_offset || !pair_cell || !workspace) return wc_fail("wc_paths_pair_order: null argument");
    if (tx_count == 0) return wc_fail("wc_paths_pair_order: cells without transmitters");
    if ((unsigned long long)max_rx * (unsigned long long)tx_count > 0xffffffffull)
        return wc_fail("wc_paths_pair_order: %d receivers x %d transmitters exceed the 32-bit pair key", max_rx,
                       tx_count);
    const CompactLayout l = compact_layout(cells);
    char* ws = static_cast<char*>(workspace);
    unsigned* keys = reinterpret_cast<unsigned*>(ws + l.keys0);
    unsigned* keys_tmp = reinterpret_cast<unsigned*>(ws + l.keys1);
    int* vals_tmp = reinterpret_cast<int*>(ws + l.vals_tmp);
    int* hist = reinterpret_cast<int*>(ws + l.hist);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(pair_keys_kernel, dim3(grid_for(cells, 256)), dim3(256), 0, st, cell_offset, tx_count, cells,
                       keys, pair_cell);
    int bits = 1;
    const unsigned long long key_max = (unsigned long long)max_rx * (unsigned long long)tx_count;
    while (bits < 32 && (key_max >> bits) != 0) ++bits;
    if (radix_sort_pairs(keys, pair_cell, keys_tmp, vals_tmp, hist, cells, bits, st) != 0)
        return wc_fail_hip("wc_paths_pair_order: radix sort", hipGetLastError());
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_paths_pair_order", err);
}

extern "C" int wc_paths_row_offsets(const int* path_count, const int* pair_cell, int cells, int* row_offset,
                                    int* pair_offset, long long* total, void* workspace, void* stream) {
    if (cells < 0) return wc_fail("wc_paths_row_offsets: cells must be non-negative");
    if (!pair_offset || !total || !workspace || (cells > 0 && (!path_count || !pair_cell || !row_offset)))
        return wc_fail("wc_paths_row_offsets: null argument");
    const CompactLayout l = compact_layout(cells);
    if (scan_launch(path_count, pair_cell, cells, pair_offset, row_offset,
                    static_cast<char*>(workspace) + l.scan, total, (hipStream_t)stream) != 0)
        return wc_fail_hip("wc_paths_row_offsets", hipGetLastError());
    return 0;
}

extern "C" int wc_paths_canonical_order(const wc_path_rows* rows, const int* pair_offset, int pair_count,
                                        int* permutation, unsigned char* keep, void* stream) {
    if (int rc = check_rows(rows, "wc_paths_canonical_order")) return rc;
    if (!pair_offset || !permutation || !keep) return wc_fail("wc_paths_canonical_order: null argument");
    if (pair_count <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(pair_rank_kernel, dim3(grid_for(pair_count, 1)), dim3(64), 0, st, *rows, pair_offset,
                       pair_count, permutation);
    hipLaunchKernelGGL(pair_keep_kernel, dim3(grid_for(pair_count, 1)), dim3(64), 0, st, *rows, pair_offset,
                       pair_count, permutation, keep, -1, nullptr, nullptr);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_paths_canonical_order", err);
}

extern "C" int wc_paths_compact(const wc_path_rows* rows, const int* pair_offset, int pair_count,
                                int max_paths_per_pair, int* permutation, unsigned char* keep, int* kept_offset,
                                int* source_row, long long* kept_total, int* dropped, void* workspace, void* stream) {
    if (int rc = check_rows(rows, "wc_paths_compact")) return rc;
    if (pair_count < 0) return wc_fail("wc_paths_compact: pair_count must be non-negative");
    if (!kept_offset || !kept_total || !dropped || !workspace ||
        (pair_count > 0 && (!pair_offset || !permutation || !keep || !source_row)))
        return wc_fail("wc_paths_compact: null argument");
    const CompactLayout l = compact_layout(pair_count);
    char* ws = static_cast<char*>(workspace);
    int* kept = reinterpret_cast<int*>(ws + l.kept);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, st, dropped, 1);
    if (pair_count > 0) {
        hipLaunchKernelGGL(pair_rank_kernel, dim3(grid_for(pair_count, 1)), dim3(64), 0, st, *rows, pair_offset,
                           pair_count, permutation);
        hipLaunchKernelGGL(pair_keep_kernel, dim3(grid_for(pair_count, 1)), dim3(64), 0, st, *rows, pair_offset,
                           pair_count, permutation, keep, max_paths_per_pair, kept, dropped);
    }
    if (scan_launch(kept, nullptr, pair_count, kept_offset, nullptr, ws + l.scan, kept_total, st) != 0)
        return wc_fail_hip("wc_paths_compact: scan", hipGetLastError());
    if (pair_count > 0)
        hipLaunchKernelGGL(pair_gather_kernel, dim3(grid_for(pair_count, 1)), dim3(64), 0, st, pair_offset,
                           pair_count, permutation, keep, kept_offset, source_row);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_paths_compact", err);
}
