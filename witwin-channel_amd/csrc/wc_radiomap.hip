// Fused radiomap, path export and discovery kernels behind the C ABI.
#include "wc_api_internal.h"
#include "wc_cone.h"
#include "wc_paths.h"
#include "../../include/wcmi.h"

namespace wc {

static_assert(sizeof(SceneView) == sizeof(wc_scene), "wc_scene must mirror SceneView");
static_assert(kConeFloats == WC_CONE_FLOATS, "cone record size");
static_assert(sizeof(Edge) == sizeof(wc_edge), "wc_edge must mirror Edge");

__device__ __forceinline__ int slot_of(int component) {
    switch (component) {
        case kCidLos: return 0;
        case kCidRefl: return 1;
        case kCidDiff: return 2;
        case kCidTrans: return 3;
        default: return -1;
    }
}

// Chain-validity bits of one evaluation position: chunk k (chains
// 32 k .. 32 k + 31 of the transmitter) is the word words[k * stride]
// (chunk-major, position-indexed, so a wavefront's 64 positions read and write
// 256 consecutive bytes); `touched` is the tile's chunk bitmap: bit k is set
// when some position of the tile has a set bit in chunk k. Words of untouched
// chunks are never read (and not kept zero).
struct ChainMaskView {
    const unsigned* words;
    size_t stride;
    const unsigned* touched;
};

// Next chain of [c, c1) whose bit is set (bit c - base), or c1.
__device__ __forceinline__ int next_set_chain(const ChainMaskView& m, int base, int c, int c1) {
    while (c < c1) {
        const int rel = c - base;
        const int chunk = rel >> 5;
        const unsigned tw = m.touched[chunk >> 5] >> (chunk & 31);
        if (tw == 0u) {  // no touched chunk up to the end of this bitmap word
            c = base + (((chunk >> 5) + 1) << 10);
            continue;
        }
        if (!(tw & 1u)) {  // jump to the next touched chunk
            c = base + ((chunk + __ffs(tw) - 1) << 5);
            continue;
        }
        const unsigned word = m.words[(size_t)chunk * m.stride] >> (rel & 31);
        if (word) {
            const int hit = c + __ffs(word) - 1;
            return hit < c1 ? hit : c1;
        }
        c += 32 - (rel & 31);
    }
    return c1;
}

// Reflection chains of one compile-time depth D for one cell. With a chain
// mask (from chain_visibility_kernel) only the chains whose legs are all
// clear are visited and their legs are not traced again.
template <int D, class Sink, bool kMasked>
__device__ __forceinline__ void reflect_depth(const SceneView& s, const wc_solve& q, int t, const V3<float>& tx,
                                              const V3<float>& rx, const V3<float>& tx_pol, const V3<float>& rx_pol,
                                              float sqrt_power, float freq, double k_d, int sub, int stride,
                                              Sink& sink, const ChainMaskView& chain_mask) {
    const int base = q.chain_offset[t * kMaxDepth];
    const int c0 = q.chain_offset[t * kMaxDepth + D - 1];
    const int c1 = q.chain_offset[t * kMaxDepth + D];
    for (int c = kMasked ? next_set_chain(chain_mask, base, c0, c1) : c0 + sub; c < c1;
         c = kMasked ? next_set_chain(chain_mask, base, c + 1, c1) : c + stride) {
        if (kMasked && (c - c0) % stride != sub) continue;
        const int* groups = q.chain_groups + c * kMaxDepth;
        const double* images = q.chain_images + c * kMaxDepth * 3;
        V3<double> hd[D];
        int faces[D];
        if (!solve_chain<D, !kMasked>(s, groups, images, tx, rx, hd, faces)) continue;
        PathRecord rec;
        rec.component = kCidRefl;
        rec.depth = D;
        rec.candidate = c;
        rec.edge = -1;
        V3<float> hf[D], nf[D];
        int mats[D];
#pragma unroll
        for (int b = 0; b < D; ++b) {
            hf[b] = to_f(hd[b]);
            double off;
            nf[b] = to_f(group_normal(s, groups[b], off));
            mats[b] = s.face_material[faces[b]];
            if constexpr (Sink::kNeedsRows) {
                rec.prim[b] = faces[b];
                rec.mat[b] = mats[b];
                rec.pos[b][0] = hf[b].x; rec.pos[b][1] = hf[b].y; rec.pos[b][2] = hf[b].z;
                rec.nrm[b][0] = nf[b].x; rec.nrm[b][1] = nf[b].y; rec.nrm[b][2] = nf[b].z;
            }
        }
        const V3<double> img{images[3 * (D - 1)], images[3 * (D - 1) + 1], images[3 * (D - 1) + 2]};
        rec.length_d = norm(to_d(rx) - img);
        rec.length = (float)rec.length_d;
        V3<float> arrival, departure;
        const C3<float> v = reflection_field<D>(s, tx, rx, hf, nf, mats, tx_pol, freq, k_d, rec.length_d, arrival,
                                                departure);
        finish_record(rec, v, arrival, rx_pol, sqrt_power);
        rec.dep[0] = departure.x; rec.dep[1] = departure.y; rec.dep[2] = departure.z;
        sink.path(rec);
    }
}

// Visibility kernels stage the tree's top levels in LDS (stage_top_nodes)
// and walk with them; 0 walks every node from HBM.
#ifndef WC_TOP_NODES_LDS
#define WC_TOP_NODES_LDS 1
#endif
constexpr bool kStageTop = WC_TOP_NODES_LDS != 0;

template <bool kTop = false>
__device__ __forceinline__ bool segment_clear(const SceneView& s, const V3<float>& a, const V3<float>& b, int ig0,
                                              int ig1) {
    const V3<float> d = b - a;
    const float len = norm(d);
    if (!(len > 2.0f * kSegmentBias)) return true;
    const float eps = kSegmentBias / len;
    const RayPre r = ray_setup(a.x, a.y, a.z, d.x, d.y, d.z);
    return !occluded<kTop>(s, r, eps, 1.0f - eps, ig0, ig1);
}

// Visibility of a diffraction candidate at the stationary point clamped to
// the finite edge; the receiver leg is tested first (street-level receivers
// see fewer edges than an elevated transmitter: on the scaled city the
// transmitter leg first costs 1,322 instead of 1,230 ms per step).
template <bool kTop = false>
__device__ __forceinline__ bool edge_visible(const SceneView& s, const Edge& e, const DiffGeometry& g,
                                             const V3<float>& tx, const V3<float>& rx) {
    const double tc = fmin(fmax(g.tstar, 0.0), (double)e.length);
    const V3<float> qc{e.ax + e.dx * (float)tc, e.ay + e.dy * (float)tc, e.az + e.dz * (float)tc};
    // one traversal instance (one stack) serves both legs
#pragma unroll 1
    for (int leg = 0; leg < 2; ++leg) {
        const V3<float> a = leg == 0 ? qc : tx;
        const V3<float> b = leg == 0 ? rx : qc;
        if (!segment_clear<kTop>(s, a, b, e.group0, e.group1)) return false;
    }
    return true;
}

// Passes of the fused evaluation: all components (row export), the
// non-diffraction components, or diffraction only. Splitting the radiomap
// into two launches gives each its own register budget.
enum EvalPass : int { kPassAll = 0, kPassMain = 1, kPassDiff = 2 };

// Walks the candidates of one (transmitter, cell) pair and hands each valid
// path to the sink. With stride 1 (sub 0) every candidate is visited in
// canonical order; with stride L the lane `sub` of the L lanes sharing a cell
// visits every L-th reflection chain and diffraction edge (line of sight and
// transmission belong to sub-lane 0).
template <class Sink, int kPass, bool kMasked = false>
__device__ void eval_cell(const SceneView& s, const wc_solve& q, int t, int cell, int sub, int stride, Sink& sink,
                          const unsigned* diff_mask = nullptr, int mask_words = 0,
                          const ChainMaskView& chain_mask = ChainMaskView{nullptr, 0, nullptr}) {
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    const V3<float> tx_pol = ld3(q.tx_pol + 3 * t);
    const float sqrt_power = sqrtf(fmaxf(q.tx_power[t], 0.0f));
    const int rx_index = q.rx_shared ? cell - q.cell_offset[t] : cell;
    const V3<float> rx = ld3(q.rx_pos + 3 * rx_index);
    const V3<float> rx_pol = ld3(q.rx_pol + 3 * (q.rx_pol_stride ? rx_index : 0));
    const float freq = (float)q.frequency_hz;
    const double k_d = q.frequency_hz * kTwoPi / kC0;
    const float k = (float)k_d;
    const int comps = q.components;
    const int max_depth = q.max_depth;
    constexpr bool kMain = kPass != kPassDiff;
    constexpr bool kDiff = kPass != kPassMain;

    // --- straight segment: line of sight and ordered wall hits
    Hit hits[kMaxHits];
    int nh = 0;
    const V3<float> seg = rx - tx;
    const double dist_d = norm(to_d(rx) - to_d(tx));
    const V3<float> dir = safe_normalize(seg, v3c<float>(0.f, 0.f, 1.f));
    if (kMain && sub == 0 && (comps & (kCompLos | kCompTrans))) {
        const float len = (float)dist_d;
        if (len > 2.0f * kSegmentBias) {
            const float eps = kSegmentBias / len;
            const RayPre r = ray_setup(tx.x, tx.y, tx.z, seg.x, seg.y, seg.z);
            if (comps & kCompTrans) {
                nh = gather_hits(s, r, eps, 1.0f - eps, max_depth, hits);
            } else {
                nh = occluded(s, r, eps, 1.0f - eps, -1, -1) ? 1 : 0;
            }
        }
    }
    if (kMain && sub == 0 && (comps & kCompLos) && nh == 0) {
        PathRecord rec;
        rec.component = kCidLos;
        rec.depth = 0;
        rec.candidate = -1;
        rec.edge = -1;
        rec.length_d = dist_d;
        rec.length = (float)dist_d;
        const C3<float> v = los_field(dir, tx_pol, k, k_d, dist_d);
        finish_record(rec, v, dir, rx_pol, sqrt_power);
        rec.dep[0] = dir.x; rec.dep[1] = dir.y; rec.dep[2] = dir.z;
        sink.path(rec);
    }
    if (kMain && sub == 0 && (comps & kCompTrans) && (comps & kStrictPenetration) && nh > max_depth)
        sink.overflow();
    for (int d = 1; d <= max_depth; ++d) {
        if (kMain && (comps & kCompRefl)) {
            switch (d) {
                case 1: reflect_depth<1, Sink, kMasked>(s, q, t, tx, rx, tx_pol, rx_pol, sqrt_power, freq, k_d, sub, stride, sink, chain_mask); break;
                case 2: reflect_depth<2, Sink, kMasked>(s, q, t, tx, rx, tx_pol, rx_pol, sqrt_power, freq, k_d, sub, stride, sink, chain_mask); break;
                case 3: reflect_depth<3, Sink, kMasked>(s, q, t, tx, rx, tx_pol, rx_pol, sqrt_power, freq, k_d, sub, stride, sink, chain_mask); break;
                case 4: reflect_depth<4, Sink, kMasked>(s, q, t, tx, rx, tx_pol, rx_pol, sqrt_power, freq, k_d, sub, stride, sink, chain_mask); break;
                default: reflect_depth<5, Sink, kMasked>(s, q, t, tx, rx, tx_pol, rx_pol, sqrt_power, freq, k_d, sub, stride, sink, chain_mask); break;
            }
        }
        if (kDiff && d == 1 && (comps & kCompDiff)) {
            // with a mask (from diff_visibility_kernel) only the visible edges
            // are visited, in ascending order; without one every edge is
            // tested here
            int word = 0;
            unsigned bits = 0;
            int ei = diff_mask ? -1 : sub - stride;
            while (true) {
                if (diff_mask) {
                    while (bits == 0 && word < mask_words) bits = diff_mask[word++];
                    if (bits == 0) break;
                    ei = (word - 1) * 32 + __ffs(bits) - 1;
                    bits &= bits - 1;
                    if (ei % stride != sub) continue;
                } else {
                    ei += stride;
                    if (ei >= s.edge_count) break;
                }
                const Edge e = s.edges[ei];
                DiffGeometry g;
                if (!diffraction_geometry(e, tx, rx, g)) continue;
                if (!diff_mask && !edge_visible(s, e, g, tx, rx)) continue;
                PathRecord rec;
                rec.component = kCidDiff;
                rec.depth = 1;
                rec.candidate = ei;
                rec.edge = ei;
                if constexpr (Sink::kNeedsRows) {
                    rec.prim[0] = ei;
                    rec.mat[0] = e.material0;
                    rec.pos[0][0] = (float)g.q.x; rec.pos[0][1] = (float)g.q.y; rec.pos[0][2] = (float)g.q.z;
                    rec.nrm[0][0] = e.n0x; rec.nrm[0][1] = e.n0y; rec.nrm[0][2] = e.n0z;
                }
                rec.length_d = g.s + g.sp;
                rec.length = (float)rec.length_d;
                V3<float> arrival, departure;
                const C3<float> v = diffraction_field(s, e, g, tx, rx, tx_pol, freq, k_d, arrival, departure);
                finish_record(rec, v, arrival, rx_pol, sqrt_power);
                if (!(rec.gain >= kGainFloor)) continue;
                rec.dep[0] = departure.x; rec.dep[1] = departure.y; rec.dep[2] = departure.z;
                sink.path(rec);
            }
        }
        if (kMain && sub == 0 && (comps & kCompTrans) && nh == d) {
            PathRecord rec;
            rec.component = kCidTrans;
            rec.depth = d;
            rec.candidate = -1;
            rec.edge = -1;
            V3<float> normals[kMaxHits];
            const C3<float> v = transmission_field(s, d, hits, dir, tx_pol, freq, k_d, dist_d, rec.mat, normals);
            if constexpr (Sink::kNeedsRows) {
                for (int w = 0; w < d; ++w) {
                    rec.prim[w] = hits[w].face;
                    const float tt = hits[w].t;
                    rec.pos[w][0] = tx.x + seg.x * tt; rec.pos[w][1] = tx.y + seg.y * tt;
                    rec.pos[w][2] = tx.z + seg.z * tt;
                    rec.nrm[w][0] = normals[w].x; rec.nrm[w][1] = normals[w].y; rec.nrm[w][2] = normals[w].z;
                }
            }
            rec.length_d = dist_d;
            rec.length = (float)dist_d;
            finish_record(rec, v, dir, rx_pol, sqrt_power);
            rec.dep[0] = dir.x; rec.dep[1] = dir.y; rec.dep[2] = dir.z;
            sink.path(rec);
        }
    }
}

// ------------------------------------------------------------------ sinks
struct AccumSink {
    static constexpr bool kNeedsRows = false;
    double s_re = 0.0, s_im = 0.0, p_inc = 0.0;
    double c_re[4] = {0, 0, 0, 0}, c_im[4] = {0, 0, 0, 0}, c_pow[4] = {0, 0, 0, 0};
    double dv[6] = {0, 0, 0, 0, 0, 0};
    int count = 0;
    int* topo = nullptr;
    int cap = 0;
    float sqrt_power = 1.0f;
    int* failure = nullptr;
    __device__ void overflow() {
        if (failure) atomicOr(failure, kFailPenetration);
    }
    __device__ void path(const PathRecord& r) {
        const int slot = slot_of(r.component);
        const double re = r.path_field.re, im = r.path_field.im;
        c_re[slot] += re;
        c_im[slot] += im;
        c_pow[slot] += (double)r.gain;
        s_re += re;
        s_im += im;
        p_inc += (double)r.gain;
        if (r.component == kCidDiff) {
            dv[0] += (double)r.field.x.re * sqrt_power; dv[1] += (double)r.field.x.im * sqrt_power;
            dv[2] += (double)r.field.y.re * sqrt_power; dv[3] += (double)r.field.y.im * sqrt_power;
            dv[4] += (double)r.field.z.re * sqrt_power; dv[5] += (double)r.field.z.im * sqrt_power;
        }
        if (topo && count < cap) topo[count] = (r.component << 24) | (r.candidate & 0xFFFFFF);
        ++count;
    }
};

struct CountSink {
    __device__ void overflow() {}
    static constexpr bool kNeedsRows = false;
    int count = 0;
    __device__ void path(const PathRecord&) { ++count; }
};

// Per-cell state carried from the main pass to the diffraction pass.
struct CellPartial {
    double s_re, s_im, p_inc;
    int count;
    int pad;
};
static_assert(sizeof(CellPartial) == 32, "CellPartial layout");

struct EmitSink {
    __device__ void overflow() {}
    static constexpr bool kNeedsRows = true;
    const wc_path_rows* rows;
    int row;
    int row_end;  // rows reserved for this cell by the count pass; never written past
    int tx, rx;
    __device__ void path(const PathRecord& r) {
        if (row >= row_end) return;
        const int k = row++;
        const int w = rows->width;
        rows->tx_id[k] = tx;
        rows->rx_id[k] = rx;
        rows->depth[k] = r.depth;
        rows->component_id[k] = r.component;
        rows->edge_id[k] = r.edge;
        rows->path_length_m[k] = r.length;
        rows->delay_s[k] = (float)(r.length_d / kC0);
        rows->path_gain[k] = r.gain;
        const bool has_seq = r.component != kCidLos;
        rows->primitive_id[k] = has_seq ? r.prim[0] : -1;
        rows->material_id[k] = has_seq ? r.mat[0] : -1;
        for (int j = 0; j < w; ++j) {
            const bool on = has_seq && j < r.depth;
            rows->primitive_sequence[k * w + j] = on ? r.prim[j] : -1;
            rows->material_sequence[k * w + j] = on ? r.mat[j] : -1;
            for (int c = 0; c < 3; ++c) {
                rows->interaction_positions[(k * w + j) * 3 + c] = on ? r.pos[j][c] : 0.0f;
                rows->interaction_normals[(k * w + j) * 3 + c] = on ? r.nrm[j][c] : 0.0f;
            }
        }
        rows->path_field[2 * k] = r.path_field.re;
        rows->path_field[2 * k + 1] = r.path_field.im;
        rows->coefficient[2 * k] = r.coefficient.re;
        rows->coefficient[2 * k + 1] = r.coefficient.im;
        float* fx = rows->field_xyz + 6 * k;
        fx[0] = r.field.x.re; fx[1] = r.field.x.im;
        fx[2] = r.field.y.re; fx[3] = r.field.y.im;
        fx[4] = r.field.z.re; fx[5] = r.field.z.im;
        for (int c = 0; c < 3; ++c) {
            rows->arrival_direction[3 * k + c] = r.arr[c];
            rows->departure_direction[3 * k + c] = r.dep[c];
        }
    }
};

__device__ __forceinline__ int tx_of_cell(const wc_solve& q, int cell) {
    int lo = 0, hi = q.tx_count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (q.cell_offset[mid] <= cell) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Workspace of a radiomap with diffraction: per-(cell, lane) partial sums
// (32 B each) carried from the reflection pass, then the per-cell edge
// visibility masks of the diffraction pass (mask_words(E) words per cell).
__host__ __device__ __forceinline__ int mask_words(int edge_count) { return (edge_count + 31) >> 5; }
__host__ __device__ __forceinline__ size_t diff_mask_bytes(int cells, int edge_count) {
    return (size_t)cells * (size_t)mask_words(edge_count) * 4;
}
__host__ __device__ __forceinline__ size_t partial_bytes(int cells, int lanes) {
    return (size_t)cells * (size_t)lanes * 32;
}
__device__ __forceinline__ unsigned* diff_masks(const wc_radiomap_out& o, const wc_solve& q) {
    return reinterpret_cast<unsigned*>(static_cast<char*>(o.workspace) + partial_bytes(q.total_cells,
                                                                                        q.lanes_per_cell));
}
// after the diffraction masks: mask_words(max_tx_chains) words per cell, bit
// c - chain_offset[t * 5] for chain c of the cell's transmitter t
__host__ __device__ __forceinline__ unsigned* chain_masks(void* workspace, const wc_solve& q, int edge_count) {
    return reinterpret_cast<unsigned*>(static_cast<char*>(workspace) + partial_bytes(q.total_cells,
                                                                                     q.lanes_per_cell) +
                                       diff_mask_bytes(q.total_cells, edge_count));
}

// after the chain masks (8-byte aligned): the transmitter-leg shadow bins of
// every (transmitter, edge), one 64-bit word each (edge_shadow_kernel)
__host__ __device__ __forceinline__ size_t chain_mask_bytes(int cells, int max_tx_chains) {
    return (size_t)cells * (size_t)mask_words(max_tx_chains > 0 ? max_tx_chains : 0) * 4;
}
// after the chain masks: one chunk bitmap per tile (touched_words(max chains)
// words each, see ChainMaskView), for at most ceil(cells * L / 64) + T tiles;
// the shadow table then takes tx_count as it is passed here
__host__ __device__ __forceinline__ int touched_words(int max_tx_chains) {
    return (mask_words(max_tx_chains > 0 ? max_tx_chains : 0) + 31) >> 5;
}
__host__ __device__ __forceinline__ size_t tile_capacity(int cells, int lanes, int tx_count) {
    return ((size_t)(cells > 0 ? cells : 0) * (size_t)lanes + 63) / 64 + (size_t)(tx_count > 0 ? tx_count : 0);
}
__host__ __device__ __forceinline__ size_t touched_offset(int cells, int lanes, int edge_count, int max_tx_chains) {
    return partial_bytes(cells, lanes) + diff_mask_bytes(cells, edge_count) + chain_mask_bytes(cells, max_tx_chains);
}
__host__ __device__ __forceinline__ unsigned* touched_bitmaps(void* workspace, const wc_solve& q, int edge_count) {
    return reinterpret_cast<unsigned*>(static_cast<char*>(workspace) +
                                       touched_offset(q.total_cells, q.lanes_per_cell, edge_count, q.max_tx_chains));
}
__host__ __device__ __forceinline__ size_t shadow_offset(int cells, int lanes, int edge_count, int max_tx_chains,
                                                         int tx_count) {
    const size_t o = touched_offset(cells, lanes, edge_count, max_tx_chains) +
                     tile_capacity(cells, lanes, tx_count) * (size_t)touched_words(max_tx_chains) * 4;
    return (o + 7) & ~(size_t)7;
}

// after the shadow table: the staged diffraction pass (diff_stage_layout):
// per-cell pair offsets, control words and the pair records
struct DiffRecord {
    float pf_re, pf_im, gain;  // path field (x sqrt power) and gain; gain < 0: below the gain floor
    float f[6];                // unit-excitation field xyz (re, im)
    int edge;
};
static_assert(sizeof(DiffRecord) == 40, "DiffRecord layout");
constexpr long long kDiffRecordCap = 1LL << 22;  // staged pairs (4 M x 40 B)
__host__ __device__ __forceinline__ size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
__host__ __device__ __forceinline__ long long diff_record_capacity(int cells, int edge_count) {
    const long long all = (long long)(cells > 0 ? cells : 0) * (long long)(edge_count > 0 ? edge_count : 0);
    return all < kDiffRecordCap ? all : kDiffRecordCap;
}
struct DiffStageLayout {
    size_t offsets, ctrl, records, end;
};
__host__ __device__ inline DiffStageLayout diff_stage_layout(size_t base, int cells, int edge_count) {
    DiffStageLayout l;
    l.offsets = align256(base);
    l.ctrl = align256(l.offsets + ((size_t)cells + 1) * 4);
    l.records = align256(l.ctrl + 64);
    l.end = l.records + (size_t)diff_record_capacity(cells, edge_count) * sizeof(DiffRecord);
    return l;
}
__host__ __device__ inline size_t shadow_end(int cells, int lanes, int edge_count, int max_tx_chains, int tx_count) {
    return shadow_offset(cells, lanes, edge_count, max_tx_chains, tx_count) +
           (size_t)(tx_count > 0 ? tx_count : 0) * (size_t)(edge_count > 0 ? edge_count : 0) * 8;
}

// after the shadow table (256-aligned): the supertile cull of the chain
// validity pass for one transmitter with many chains (supertile = 16 tiles =
// 1,024 consecutive positions): one receiver box (6 floats) per supertile,
// then mask_words(max_tx_chains) words of surviving chains per supertile
constexpr int kSuperTiles = 16;
constexpr int kSuperCells = 64 * kSuperTiles;
struct SuperLayout {
    size_t boxes, masks, end;
};
// supertiles of one transmitter per 1,024 positions, plus one partial
// supertile per transmitter at most
__host__ __device__ inline size_t super_capacity(int cells, int tx_count) {
    return (size_t)((cells > 0 ? cells : 0) + kSuperCells - 1) / kSuperCells + (size_t)(tx_count > 0 ? tx_count : 0);
}
__host__ __device__ inline SuperLayout super_layout(size_t base, int cells, int max_tx_chains, int tx_count) {
    SuperLayout l;
    const size_t supers = super_capacity(cells, tx_count);
    l.boxes = align256(base);
    l.masks = align256(l.boxes + supers * 24);
    l.end = l.masks + supers * (size_t)mask_words(max_tx_chains > 0 ? max_tx_chains : 0) * 4;
    return l;
}
// start of the staged diffraction pass
__host__ __device__ inline size_t stage_base(int cells, int lanes, int edge_count, int max_tx_chains, int tx_count) {
    return super_layout(shadow_end(cells, lanes, edge_count, max_tx_chains, tx_count), cells, max_tx_chains,
                        tx_count).end;
}

// cell evaluated at launch position `pos` (identity without an order table)
__device__ __forceinline__ int cell_at(const wc_solve& q, int pos) {
    return q.cell_order ? q.cell_order[pos] : pos;
}

__device__ __forceinline__ void write_slot(const wc_solve& q, const wc_radiomap_out& o, int cell, int sl,
                                           const AccumSink& a) {
    const int cells = q.total_cells;
    const double cp = q.coherent ? a.c_re[sl] * a.c_re[sl] + a.c_im[sl] * a.c_im[sl] : a.c_pow[sl];
    o.component_power[sl * cells + cell] = (float)cp;
    o.component_field[2 * (sl * cells + cell)] = (float)a.c_re[sl];
    o.component_field[2 * (sl * cells + cell) + 1] = (float)a.c_im[sl];
}

__device__ __forceinline__ void write_totals(const wc_solve& q, const wc_radiomap_out& o, int cell,
                                             const AccumSink& a) {
    double pt, fre, fim;
    if (q.coherent) {
        fre = a.s_re;
        fim = a.s_im;
        pt = fre * fre + fim * fim;
    } else {
        pt = a.p_inc;
        fre = sqrt(fmax(pt, 0.0));
        fim = 0.0;
    }
    o.power_total[cell] = (float)pt;
    o.field_total[2 * cell] = (float)fre;
    o.field_total[2 * cell + 1] = (float)fim;
    if (o.diffraction_vector) {
        for (int j = 0; j < 6; ++j) o.diffraction_vector[6 * cell + j] = (float)a.dv[j];
    }
    if (o.path_count) o.path_count[cell] = a.count;
    if (o.topology && a.count > o.topology_capacity && o.failure) atomicOr(o.failure, 2);
}

// Fixed butterfly over the `width` consecutive lanes sharing a cell. Every
// pair adds (a + b) and (b + a), which are equal in IEEE arithmetic, so all
// lanes of the group hold the bit-identical total.
__device__ __forceinline__ double group_sum(double v, int width) {
    for (int off = 1; off < width; off <<= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ int group_sum_i(int v, int width) {
    for (int off = 1; off < width; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ void reduce_slot(AccumSink& a, int sl, int width) {
    a.c_re[sl] = group_sum(a.c_re[sl], width);
    a.c_im[sl] = group_sum(a.c_im[sl], width);
    a.c_pow[sl] = group_sum(a.c_pow[sl], width);
}

// One pass of the radiomap for one lane of a cell's lane group (L = width
// lanes per cell). kPassMain leaves each lane's running totals in the
// workspace when diffraction follows; the last pass reduces the group and
// sub-lane 0 writes the cell. Descriptor segment of a lane: topology[cell]
// [sub * cap / L ...], terminated by -1 when not full.
// kMasked: the tiled launch, whose reflection chains come from the chain
// masks (chain_visibility_kernel ran first); the plain launch traces them.
template <int kPass, bool kMasked>
__device__ __forceinline__ void radiomap_lane(const SceneView& s, const wc_solve& q, const wc_radiomap_out& o,
                                              int t, int cell, bool active, int sub, int width,
                                              int tile = 0, int pos = 0) {
    AccumSink a;
    a.sqrt_power = sqrtf(fmaxf(q.tx_power[t], 0.0f));
    a.failure = o.failure;
    CellPartial* partial = reinterpret_cast<CellPartial*>(o.workspace);
    const size_t lane_slot = (size_t)cell * width + sub;
    int cap_sub = 0;
    if (active && kPass == kPassDiff) {
        const CellPartial p = partial[lane_slot];
        a.s_re = p.s_re;
        a.s_im = p.s_im;
        a.p_inc = p.p_inc;
        a.count = p.count;
    }
    if (active && o.topology) {
        cap_sub = o.topology_capacity / width;
        a.topo = o.topology + (size_t)cell * o.topology_capacity + (size_t)sub * cap_sub;
        a.cap = cap_sub;
    }
    if (active) {
        if (kPass == kPassDiff) {
            const int words = mask_words(s.edge_count);
            const unsigned* mask = diff_masks(o, q) + (size_t)cell * words;
            eval_cell<AccumSink, kPass>(s, q, t, cell, sub, width, a, mask, words);
        } else {
            // without reflection chains the masked chain loops are empty
            ChainMaskView cmask{nullptr, 0, nullptr};
            if (kMasked && q.max_tx_chains > 0) {
                cmask.words = chain_masks(o.workspace, q, s.edge_count) + pos;
                cmask.stride = (size_t)q.total_cells;
                cmask.touched = touched_bitmaps(o.workspace, q, s.edge_count) +
                                (size_t)tile * touched_words(q.max_tx_chains);
            }
            eval_cell<AccumSink, kPass, kMasked>(s, q, t, cell, sub, width, a, nullptr, 0, cmask);
        }
    }
    if (active && o.topology) {
        if (a.count < cap_sub) a.topo[a.count] = -1;
        if (a.count > cap_sub && o.failure) atomicOr(o.failure, 2);
    }
    const bool diffraction = (q.components & kCompDiff) != 0;
    if (kPass == kPassMain && diffraction) {
        if (active) partial[lane_slot] = CellPartial{a.s_re, a.s_im, a.p_inc, a.count, 0};
        reduce_slot(a, 0, width);
        reduce_slot(a, 1, width);
        reduce_slot(a, 3, width);
        if (active && sub == 0) {
            write_slot(q, o, cell, 0, a);
            write_slot(q, o, cell, 1, a);
            write_slot(q, o, cell, 3, a);
        }
        return;
    }
    a.s_re = group_sum(a.s_re, width);
    a.s_im = group_sum(a.s_im, width);
    a.p_inc = group_sum(a.p_inc, width);
    for (int j = 0; j < 6; ++j) a.dv[j] = group_sum(a.dv[j], width);
    a.count = group_sum_i(a.count, width);
    if (kPass == kPassMain) {
        reduce_slot(a, 0, width);
        reduce_slot(a, 1, width);
        reduce_slot(a, 3, width);
    }
    reduce_slot(a, 2, width);
    if (!active || sub != 0) return;
    if (kPass == kPassMain) {
        write_slot(q, o, cell, 0, a);
        write_slot(q, o, cell, 1, a);
        write_slot(q, o, cell, 3, a);
    }
    write_slot(q, o, cell, 2, a);
    write_totals(q, o, cell, a);
}

// Occupancy targets (waves per SIMD) of the two radiomap passes; 1 leaves
// the register budget to the compiler.
#ifndef WC_MAIN_WAVES_PER_EU
#define WC_MAIN_WAVES_PER_EU 1
#endif
#ifndef WC_DIFF_WAVES_PER_EU
#define WC_DIFF_WAVES_PER_EU 1
#endif

template <int kPass>
__global__ void __launch_bounds__(256) radiomap_kernel(SceneView s, wc_solve q, wc_radiomap_out o) {
    // one lane per cell (L = 1)
    for (int base = blockIdx.x * blockDim.x; base < q.total_cells; base += blockDim.x * gridDim.x) {
        const int pos = base + threadIdx.x;
        const bool active = pos < q.total_cells;
        const int t = active ? tx_of_cell(q, pos) : 0;
        radiomap_lane<kPass, false>(s, q, o, t, cell_at(q, active ? pos : 0), active, 0, 1);
    }
}

// Persistent variant: wavefronts pull tiles (tx, first cell) of 64 / L cells
// from a global counter; each cell is evaluated by L consecutive lanes that
// split its reflection chains and diffraction edges. Tiles never mix
// transmitters, and uneven receiver sets balance across the machine.
template <int kPass>
__device__ __forceinline__ void radiomap_tiles(const SceneView& s, const wc_solve& q, const wc_radiomap_out& o,
                                               const int* __restrict__ tiles, int tile_count,
                                               int* __restrict__ counter) {
    const int lane = threadIdx.x & 63;
    const int width = q.lanes_per_cell;
    const int sub = lane & (width - 1);
    const int local = lane / width;
    while (true) {
        int tile = 0;
        if (lane == 0) tile = atomicAdd(counter, 1);
        tile = __shfl(tile, 0);
        if (tile >= tile_count) break;
        const int t = tiles[2 * tile];
        const int pos = tiles[2 * tile + 1] + local;
        const bool active = pos < q.cell_offset[t + 1];
        radiomap_lane<kPass, true>(s, q, o, t, cell_at(q, active ? pos : q.cell_offset[t]), active, sub, width, tile,
                                   pos);
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WC_MAIN_WAVES_PER_EU)))
radiomap_main_tiles(SceneView s, wc_solve q, wc_radiomap_out o, const int* __restrict__ tiles, int tile_count,
                    int* __restrict__ counter) {
    radiomap_tiles<kPassMain>(s, q, o, tiles, tile_count, counter);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WC_DIFF_WAVES_PER_EU)))
radiomap_diff_tiles(SceneView s, wc_solve q, wc_radiomap_out o, const int* __restrict__ tiles, int tile_count,
                    int* __restrict__ counter, const int* __restrict__ staged) {
    if (staged && staged[2]) return;  // the staged pass below covered every pair
    radiomap_tiles<kPassDiff>(s, q, o, tiles, tile_count, counter);
}

// ---- staged diffraction field pass ------------------------------------------
// With one lane per cell, the field pass lasts as long as the busiest cell
// (a few hundred visible edges near the transmitter against a mean of ~15),
// and a tile of such cells cannot be shared between wavefronts without
// changing the order of its float64 sums. The staged pass separates the two:
//   1. diff_count_kernel + diff_scan_kernel: visible pairs per cell and their
//      exclusive prefix over cells (one workgroup);
//   2. diff_record_kernel: every (cell, visible edge) pair evaluated once,
//      the pairs dealt in chunks of kDiffChunk over all wavefronts (the same
//      geometry, field and gain-floor test as the per-cell pass), one
//      40-byte record per pair;
//   3. diff_sum_kernel: one lane per cell sums its records in ascending edge
//      order into the carried float64 totals and appends the descriptors,
//      exactly as the per-cell pass does, so the outputs are bit-identical.
// When the pairs exceed the record capacity the control word says so, these
// kernels return at once and the per-cell pass (radiomap_diff_tiles) runs.
constexpr int kDiffChunk = 256;

__global__ void __launch_bounds__(256) diff_count_kernel(wc_solve q, const unsigned* __restrict__ masks, int words,
                                                         int* __restrict__ offsets) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= q.total_cells) return;
    int n = 0;
    for (int w = 0; w < words; ++w) n += __popc(masks[(size_t)cell * words + w]);
    offsets[cell] = n;
}

// In-place exclusive scan of offsets[0..n) by one workgroup of 1024 threads,
// 4096 consecutive entries per pass (4 per thread, coalesced): each thread's
// 4, a scan inside each wavefront, a scan of the 16 wavefront totals in LDS,
// and a running carry across passes. offsets[n] = total; ctrl[0] = total,
// ctrl[1] = 0 (chunk counter), ctrl[2] = 1 when the records fit.
__global__ void __launch_bounds__(1024) diff_scan_kernel(int* __restrict__ offsets, int n, long long capacity,
                                                         int* __restrict__ ctrl) {
    __shared__ int wave_total[16];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    long long carry = 0;  // the same in every thread
    for (int base = 0; base < n; base += 4096) {
        const int i0 = base + 4 * t;
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = i0 + k < n ? offsets[i0 + k] : 0;
        const int mine = v[0] + v[1] + v[2] + v[3];
        int x = mine;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wave_total[wid] = x;
        __syncthreads();
        if (wid == 0) {
            int w = lane < 16 ? wave_total[lane] : 0;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const int y = __shfl_up(w, off);
                if (lane >= off) w += y;
            }
            if (lane < 16) wave_total[lane] = w;
        }
        __syncthreads();
        long long run = carry + (wid > 0 ? wave_total[wid - 1] : 0) + x - mine;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < n) offsets[i0 + k] = (int)run;
            run += v[k];
        }
        carry += wave_total[15];
        __syncthreads();  // wave_total is rewritten by the next pass
    }
    if (t == 0) {
        const bool fits = carry <= capacity && carry < 0x7fffffffLL;
        offsets[n] = fits ? (int)carry : 0;
        ctrl[0] = fits ? (int)carry : 0;
        ctrl[1] = 0;
        ctrl[2] = fits ? 1 : 0;
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WC_DIFF_WAVES_PER_EU)))
diff_record_kernel(SceneView s, wc_solve q, const unsigned* __restrict__ masks, int words,
                   const int* __restrict__ offsets, int* __restrict__ ctrl, DiffRecord* __restrict__ records) {
    const int lane = threadIdx.x & 63;
    if (!ctrl[2]) return;
    const int total = ctrl[0];
    const int cells = q.total_cells;
    const float freq = (float)q.frequency_hz;
    const double k_d = q.frequency_hz * kTwoPi / kC0;
    while (true) {
        int chunk = 0;
        if (lane == 0) chunk = atomicAdd(ctrl + 1, 1);
        chunk = __shfl(chunk, 0);
        const long long base = (long long)chunk * kDiffChunk;
        if (base >= total) break;
#pragma unroll 1
        for (int j = 0; j < kDiffChunk / 64; ++j) {
            const long long p64 = base + j * 64 + lane;
            if (p64 >= total) break;
            const int p = (int)p64;
            // owner cell: the last cell whose offset is <= p
            int lo = 0, hi = cells - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (offsets[mid] <= p) lo = mid;
                else hi = mid - 1;
            }
            const int cell = lo;
            int k = p - offsets[cell];
            const unsigned* m = masks + (size_t)cell * words;
            int w = 0;
            unsigned bits = m[0];
            while (k >= __popc(bits)) {
                k -= __popc(bits);
                bits = m[++w];
            }
            for (int i = 0; i < k; ++i) bits &= bits - 1;
            const int ei = w * 32 + __ffs(bits) - 1;
            const int t = tx_of_cell(q, cell);
            const V3<float> tx = ld3(q.tx_pos + 3 * t);
            const V3<float> tx_pol = ld3(q.tx_pol + 3 * t);
            const float sqrt_power = sqrtf(fmaxf(q.tx_power[t], 0.0f));
            const int rx_index = q.rx_shared ? cell - q.cell_offset[t] : cell;
            const V3<float> rx = ld3(q.rx_pos + 3 * rx_index);
            const V3<float> rx_pol = ld3(q.rx_pol + 3 * (q.rx_pol_stride ? rx_index : 0));
            const Edge e = s.edges[ei];
            DiffRecord r;
            r.edge = ei;
            r.pf_re = r.pf_im = 0.0f;
            r.gain = -1.0f;
            for (int i = 0; i < 6; ++i) r.f[i] = 0.0f;
            DiffGeometry g;
            if (diffraction_geometry(e, tx, rx, g)) {
                V3<float> arrival, departure;
                const C3<float> v = diffraction_field(s, e, g, tx, rx, tx_pol, freq, k_d, arrival, departure);
                PathRecord rec;
                finish_record(rec, v, arrival, rx_pol, sqrt_power);
                if (rec.gain >= kGainFloor) {
                    r.pf_re = rec.path_field.re;
                    r.pf_im = rec.path_field.im;
                    r.gain = rec.gain;
                    r.f[0] = v.x.re; r.f[1] = v.x.im;
                    r.f[2] = v.y.re; r.f[3] = v.y.im;
                    r.f[4] = v.z.re; r.f[5] = v.z.im;
                }
            }
            records[p] = r;
        }
    }
}

__global__ void __launch_bounds__(256) diff_sum_kernel(wc_solve q, wc_radiomap_out o, const int* __restrict__ offsets,
                                                       const int* __restrict__ ctrl,
                                                       const DiffRecord* __restrict__ records) {
    // evaluation order: neighbouring receivers (similar record counts) share a wavefront
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= q.total_cells || !ctrl[2]) return;
    const int cell = cell_at(q, pos);
    const int t = tx_of_cell(q, pos);
    AccumSink a;
    a.sqrt_power = sqrtf(fmaxf(q.tx_power[t], 0.0f));
    a.failure = o.failure;
    const CellPartial p = reinterpret_cast<const CellPartial*>(o.workspace)[cell];
    a.s_re = p.s_re;
    a.s_im = p.s_im;
    a.p_inc = p.p_inc;
    a.count = p.count;
    int cap = 0;
    if (o.topology) {
        cap = o.topology_capacity;
        a.topo = o.topology + (size_t)cell * cap;
        a.cap = cap;
    }
    for (int i = offsets[cell]; i < offsets[cell + 1]; ++i) {
        const DiffRecord r = records[i];
        if (!(r.gain >= 0.0f)) continue;
        PathRecord rec;
        rec.component = kCidDiff;
        rec.candidate = r.edge;
        rec.path_field = {r.pf_re, r.pf_im};
        rec.gain = r.gain;
        rec.field.x = {r.f[0], r.f[1]};
        rec.field.y = {r.f[2], r.f[3]};
        rec.field.z = {r.f[4], r.f[5]};
        a.path(rec);
    }
    if (o.topology) {
        if (a.count < cap) a.topo[a.count] = -1;
        if (a.count > cap && o.failure) atomicOr(o.failure, 2);
    }
    write_slot(q, o, cell, 2, a);
    write_totals(q, o, cell, a);
}

// Chain validity for every (cell, reflection chain) candidate: bit c - base
// of the cell's chain mask is set when the back-projection through the
// chain's images lands inside a face of every group and all D + 1 legs are
// clear. One wavefront per (tile of one transmitter, 32 consecutive chains of
// that transmitter; the chain and its depth are wave-uniform). The wavefront
// first culls its 32 chains against the bounding box of the tile's receivers
// (chain_reaches_box, one chain per lane), then every receiver tests the
// surviving chains exactly. Only geometry and traversal live here, so the
// kernel keeps many wavefronts per SIMD in flight; the reflection field pass
// then visits the valid chains only.
template <int D, bool kTop = false>
__device__ __forceinline__ bool chain_valid(const SceneView& s, const wc_solve& q, int c, const V3<float>& tx,
                                            const V3<float>& rx) {
    V3<double> hd[D];
    int faces[D];
    return solve_chain<D, true, kTop>(s, q.chain_groups + c * kMaxDepth, q.chain_images + c * kMaxDepth * 3, tx, rx, hd,
                                faces);
}

// Conservative tile test of one chain: can the last bounce of chain c (depth
// D) land inside its group for some receiver in the box [lo, hi]? The hit of
// the last leg is the central projection of the receiver from the last image
// onto the group plane, so over the box it stays inside the convex hull of the
// projected box corners when every corner lies across the plane from the
// image; with no corner across, no receiver of the box can reach the plane.
// false only when the chain fails solve_chain's last-bounce tests for every
// receiver of the box (margin 1e-3 m plus 1e-5 of the coordinates).
__device__ inline bool chain_reaches_box(const SceneView& s, const wc_solve& q, int c, int D, const V3<float>& lo,
                                         const V3<float>& hi) {
    // the chain's receiver cone first (nested bounces, convex group hulls)
    if (q.chain_cones && !cone_reaches_box(q.chain_cones + (size_t)c * kConeFloats, lo, hi)) return false;
    const int g = q.chain_groups[c * kMaxDepth + D - 1];
    const double* im = q.chain_images + ((size_t)c * kMaxDepth + D - 1) * 3;
    double off;
    const V3<double> n = group_normal(s, g, off);
    const V3<double> I{im[0], im[1], im[2]};
    const double si = dot(n, I) - off;
    V3<double> x[8];
    int across = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const V3<double> corner{(k & 1) ? hi.x : lo.x, (k & 2) ? hi.y : lo.y, (k & 4) ? hi.z : lo.z};
        const double sn = dot(n, corner) - off;
        if (si * sn < 0.0) ++across;
        x[k] = I + (corner - I) * (si / (si - sn));
    }
    if (across == 0) return false;
    if (across < 8) return true;
    V3<double> plo = x[0], phi = x[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        plo = {fmin(plo.x, x[k].x), fmin(plo.y, x[k].y), fmin(plo.z, x[k].z)};
        phi = {fmax(phi.x, x[k].x), fmax(phi.y, x[k].y), fmax(phi.z, x[k].z)};
    }
    // the hull of the projected corners misses a face when their boxes are
    // apart or all corners lie outside one of its edges; the chain misses the
    // box when that holds for every face of the group
    for (int i = s.group_face_offset[g]; i < s.group_face_offset[g + 1]; ++i) {
        const Tri tri = s.faces[s.group_faces[i]];
        const V3<double> v[3] = {{tri.v0.x, tri.v0.y, tri.v0.z}, {tri.v1.x, tri.v1.y, tri.v1.z},
                                 {tri.v2.x, tri.v2.y, tri.v2.z}};
        const V3<double> flo{fmin(v[0].x, fmin(v[1].x, v[2].x)), fmin(v[0].y, fmin(v[1].y, v[2].y)),
                             fmin(v[0].z, fmin(v[1].z, v[2].z))};
        const V3<double> fhi{fmax(v[0].x, fmax(v[1].x, v[2].x)), fmax(v[0].y, fmax(v[1].y, v[2].y)),
                             fmax(v[0].z, fmax(v[1].z, v[2].z))};
        const double mb = 1.0e-3 + 1.0e-5 * (fabs(flo.x) + fabs(flo.y) + fabs(flo.z) + fabs(fhi.x) + fabs(fhi.y) +
                                             fabs(fhi.z));
        if (plo.x > fhi.x + mb || phi.x < flo.x - mb || plo.y > fhi.y + mb || phi.y < flo.y - mb ||
            plo.z > fhi.z + mb || phi.z < flo.z - mb)
            continue;
        const double orient = dot(cross(v[1] - v[0], v[2] - v[0]), n) >= 0.0 ? 1.0 : -1.0;
        bool separated = false;
        for (int e = 0; e < 3 && !separated; ++e) {
            const V3<double> a = v[e], b = v[(e + 1) % 3];
            const V3<double> ab = b - a;
            const double m = (1.0e-3 + 1.0e-5 * (fabs(a.x) + fabs(a.y) + fabs(a.z))) * norm(ab);
            bool all_out = true;
#pragma unroll
            for (int k = 0; k < 8; ++k) all_out = all_out && orient * dot(cross(ab, x[k] - a), n) < -m;
            separated = all_out;
        }
        if (!separated) return true;
    }
    return false;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

// Supertile cull (one transmitter, positions [1024 k, 1024 k + 1024)): the
// receiver box of each supertile, then every chain against each box, so the
// tile pass culls only the chains its supertile kept (a chain that misses the
// supertile box misses every tile box inside it: chain_reaches_box only
// grows with the box).
// Supertile s: its transmitter and its position range. With host supertile
// tables (super_first_tile) a super is up to 16 consecutive tiles of one
// transmitter; without them (one transmitter) it is positions
// [1024 s, 1024 s + 1024).
__device__ __forceinline__ void super_range(const wc_solve& q, const int* tiles, int sup, int& t, int& first,
                                            int& last) {
    if (q.super_first_tile) {
        const int f = q.super_first_tile[sup], l = q.super_first_tile[sup + 1];
        t = tiles[2 * f];
        first = tiles[2 * f + 1];
        last = min(tiles[2 * (l - 1) + 1] + 64, q.cell_offset[t + 1]);
    } else {
        t = 0;
        first = sup * kSuperCells;
        last = min(first + kSuperCells, q.total_cells);
    }
}

__global__ void __launch_bounds__(256) super_box_kernel(wc_solve q, const int* __restrict__ tiles, int supers,
                                                        float* __restrict__ boxes) {
    const int sup = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (sup >= supers) return;
    const int lane = threadIdx.x & 63;
    int t, first, last;
    super_range(q, tiles, sup, t, first, last);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int pos = first + lane; pos < last; pos += 64) {
        const int cell = cell_at(q, pos);
        const V3<float> rx = ld3(q.rx_pos + 3 * (q.rx_shared ? cell - q.cell_offset[t] : cell));
        lo[0] = fminf(lo[0], rx.x); lo[1] = fminf(lo[1], rx.y); lo[2] = fminf(lo[2], rx.z);
        hi[0] = fmaxf(hi[0], rx.x); hi[1] = fmaxf(hi[1], rx.y); hi[2] = fmaxf(hi[2], rx.z);
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            boxes[6 * sup + a] = lo[a];
            boxes[6 * sup + 3 + a] = hi[a];
        }
}

__global__ void __launch_bounds__(256) super_cull_kernel(SceneView s, wc_solve q, const int* __restrict__ tiles,
                                                         int supers, int words, const float* __restrict__ boxes,
                                                         unsigned* __restrict__ super_masks, long long wave_base) {
    const long long wave = wave_base + (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= (long long)supers * words) return;
    const int sup = (int)(wave / words);
    const int chunk = (int)(wave - (long long)sup * words);
    const int lane = threadIdx.x & 63;
    int t, first, last;
    super_range(q, tiles, sup, t, first, last);
    const V3<float> lo{boxes[6 * sup], boxes[6 * sup + 1], boxes[6 * sup + 2]};
    const V3<float> hi{boxes[6 * sup + 3], boxes[6 * sup + 4], boxes[6 * sup + 5]};
    const int base = q.chain_offset[t * kMaxDepth];
    const int end = q.chain_offset[t * kMaxDepth + min(q.max_depth, kMaxDepth)];
    const int c0 = base + chunk * 32;
    const int c1 = min(c0 + 32, end);
    bool reach = false;
    if (lane < c1 - c0 && first < last) {
        const int c = c0 + lane;
        int d = 1;
        while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
        reach = chain_reaches_box(s, q, c, d, lo, hi);
    }
    const unsigned bits = (unsigned)__ballot(reach);
    if (lane == 0) super_masks[(size_t)sup * words + chunk] = bits;
}

// supertile of a tile (see super_range)
__device__ __forceinline__ int super_of_tile(const wc_solve& q, int tile) {
    return q.tile_super ? q.tile_super[tile] : tile / kSuperTiles;
}

__global__ void __launch_bounds__(256) chain_visibility_kernel(SceneView s, wc_solve q, const int* __restrict__ tiles,
                                                               int tile_count, unsigned* __restrict__ masks,
                                                               int words, const unsigned* __restrict__ super_masks, long long wave_base,
                                                               unsigned* __restrict__ touched) {
    if (kStageTop) stage_top_nodes(s);
    const long long wave = wave_base + (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= (long long)tile_count * words) return;
    const int tile = (int)(wave / words);
    const int chunk = (int)(wave - (long long)tile * words);
    const int t = tiles[2 * tile];
    const int width = q.lanes_per_cell;
    const int lane = threadIdx.x & 63;
    const int pos = tiles[2 * tile + 1] + lane;
    const bool live = lane < 64 / width && pos < q.cell_offset[t + 1];
    const int cell = live ? cell_at(q, pos) : 0;
    // chains the tile's supertile kept (all when there is no supertile cull;
    // with it the masks were zero-filled first, so a tile whose supertile kept
    // none of the 32 chains has nothing to write)
    const unsigned kept = super_masks ? super_masks[(size_t)super_of_tile(q, tile) * words + chunk] : ~0u;
    if (kept == 0u) return;
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    V3<float> rx{0.0f, 0.0f, 0.0f};
    if (live) rx = ld3(q.rx_pos + 3 * (q.rx_shared ? cell - q.cell_offset[t] : cell));
    // receiver box of the tile (the whole wave takes part)
    const V3<float> lo{wave_min(live ? rx.x : INFINITY), wave_min(live ? rx.y : INFINITY),
                       wave_min(live ? rx.z : INFINITY)};
    const V3<float> hi{wave_max(live ? rx.x : -INFINITY), wave_max(live ? rx.y : -INFINITY),
                       wave_max(live ? rx.z : -INFINITY)};
    const int base = q.chain_offset[t * kMaxDepth];
    const int end = q.chain_offset[t * kMaxDepth + min(q.max_depth, kMaxDepth)];
    const int c0 = base + chunk * 32;
    const int c1 = min(c0 + 32, end);
    // lane l < 32 culls chain c0 + l against the box
    bool reach = false;
    if (lane < c1 - c0 && ((kept >> lane) & 1u)) {
        const int c = c0 + lane;
        int d = 1;
        while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
        reach = chain_reaches_box(s, q, c, d, lo, hi);
    }
    unsigned todo = (unsigned)__ballot(reach);
    unsigned bits = 0;
    int d = 1;
    if (!live) todo = 0u;
    while (todo) {
        const int b = __ffs(todo) - 1;
        todo &= todo - 1;
        const int c = c0 + b;
        while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
        bool ok;
        switch (d) {
            case 1: ok = chain_valid<1, kStageTop>(s, q, c, tx, rx); break;
            case 2: ok = chain_valid<2, kStageTop>(s, q, c, tx, rx); break;
            case 3: ok = chain_valid<3, kStageTop>(s, q, c, tx, rx); break;
            case 4: ok = chain_valid<4, kStageTop>(s, q, c, tx, rx); break;
            default: ok = chain_valid<5, kStageTop>(s, q, c, tx, rx); break;
        }
        if (ok) bits |= 1u << b;
    }
    if (live) masks[(size_t)chunk * q.total_cells + pos] = bits;
    if (__ballot(bits != 0u) != 0ull && lane == 0)
        atomicOr(&touched[(size_t)tile * ((words + 31) >> 5) + (chunk >> 5)], 1u << (chunk & 31));
}

template <int D>
__device__ __forceinline__ bool chain_lands(const SceneView& s, const wc_solve& q, int c, const V3<float>& tx,
                                            const V3<float>& rx);
template <int D, bool kTop = false>
__device__ __forceinline__ bool chain_legs(const SceneView& s, const wc_solve& q, int c, const V3<float>& tx,
                                           const V3<float>& rx);

// Chain validity behind a supertile cull, tile loops: one wavefront per
// (tile, segment of 64 chain words). Its lanes load the segment's 64 supertile
// words at once and the wavefront visits only the chunks (32 chains) its
// supertile kept: the tile cull of those chains against the tile's receiver
// box, then every receiver tests the survivors exactly. The bits equal
// chain_visibility_kernel's; mask words of chunks with no surviving chain stay
// as the zero fill left them.
// Occupancy target of the segment kernel: 4 waves per SIMD (128 VGPRs, a
// few spilled) instead of the 3 its 168 registers allow: its leg walks are
// latency-bound (scaled city chain pass 61.5 -> 50.7 ms per batch; 5 waves
// gave 51.1 ms).
#ifndef WC_SEG_WAVES_PER_EU
#define WC_SEG_WAVES_PER_EU 4
#endif

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WC_SEG_WAVES_PER_EU)))
chain_visibility_seg_kernel(SceneView s, wc_solve q,
                                                                   const int* __restrict__ tiles, int tile_count,
                                                                   unsigned* __restrict__ masks, int words,
                                                                   const unsigned* __restrict__ super_masks,
                                                                   long long wave_base,
                                                                   unsigned* __restrict__ touched) {
    if (kStageTop) stage_top_nodes(s);
    const int segs = (words + 63) >> 6;
    const long long wave = wave_base + (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= (long long)tile_count * segs) return;
    const int tile = (int)(wave / segs);
    const int seg = (int)(wave - (long long)tile * segs);
    const int t = tiles[2 * tile];
    const int lane = threadIdx.x & 63;
    const int base = q.chain_offset[t * kMaxDepth];
    const int end = q.chain_offset[t * kMaxDepth + min(q.max_depth, kMaxDepth)];
    const int tx_words = mask_words(end - base);
    const int w = seg * 64 + lane;
    const unsigned sword = w < tx_words ? super_masks[(size_t)super_of_tile(q, tile) * words + w] : 0u;
    unsigned long long chunks = __ballot(sword != 0u);
    if (chunks == 0ull) return;
    const int pos = tiles[2 * tile + 1] + lane;
    const bool live = pos < q.cell_offset[t + 1];
    const int cell = live ? cell_at(q, pos) : 0;
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    V3<float> rx{0.0f, 0.0f, 0.0f};
    if (live) rx = ld3(q.rx_pos + 3 * (q.rx_shared ? cell - q.cell_offset[t] : cell));
    const V3<float> lo{wave_min(live ? rx.x : INFINITY), wave_min(live ? rx.y : INFINITY),
                       wave_min(live ? rx.z : INFINITY)};
    const V3<float> hi{wave_max(live ? rx.x : -INFINITY), wave_max(live ? rx.y : -INFINITY),
                       wave_max(live ? rx.z : -INFINITY)};
    int d = 1;
    while (chunks) {
        const int b = __ffsll((long long)chunks) - 1;
        chunks &= chunks - 1;
        const int chunk = seg * 64 + b;
        const unsigned kept = (unsigned)__shfl((int)sword, b);
        const int c0 = base + chunk * 32;
        const int c1 = min(c0 + 32, end);
        bool reach = false;
        if (lane < c1 - c0 && ((kept >> lane) & 1u)) {
            const int c = c0 + lane;
            int dd = 1;
            while (c >= q.chain_offset[t * kMaxDepth + dd]) ++dd;
            reach = chain_reaches_box(s, q, c, dd, lo, hi);
        }
        unsigned todo = (unsigned)__ballot(reach);
        if (todo == 0u) continue;
        unsigned bits = 0;
        while (todo) {
            const int bb = __ffs(todo) - 1;
            todo &= todo - 1;
            const int c = c0 + bb;
            while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
            if (!live) continue;
            bool ok;
            switch (d) {
                case 1: ok = chain_valid<1, kStageTop>(s, q, c, tx, rx); break;
                case 2: ok = chain_valid<2, kStageTop>(s, q, c, tx, rx); break;
                case 3: ok = chain_valid<3, kStageTop>(s, q, c, tx, rx); break;
                case 4: ok = chain_valid<4, kStageTop>(s, q, c, tx, rx); break;
                default: ok = chain_valid<5, kStageTop>(s, q, c, tx, rx); break;
            }
            if (ok) bits |= 1u << bb;
        }
        if (live) masks[(size_t)chunk * q.total_cells + pos] = bits;
        if (__ballot(bits != 0u) != 0ull && lane == 0)
            atomicOr(&touched[(size_t)tile * ((words + 31) >> 5) + (chunk >> 5)], 1u << (chunk & 31));
    }
}

// The same bits with the legs queued: per (tile, segment) wavefront as
// chain_visibility_seg_kernel, but each lane first back-projects its receiver
// through the surviving chains of a chunk (no traversal), and the (receiver,
// chain) pairs that land in their faces queue up across chunks; their D+1 legs
// are traced 64 pairs per round with every lane busy, and each visible pair
// sets its bit with an atomic OR into the zero-filled masks. A wavefront's
// queue lives in its own LDS rows (wave barriers only).
__global__ void __launch_bounds__(256) chain_visibility_seg_queue_kernel(SceneView s, wc_solve q,
                                                                         const int* __restrict__ tiles,
                                                                         int tile_count, unsigned* __restrict__ masks,
                                                                         int words,
                                                                         const unsigned* __restrict__ super_masks,
                                                                         long long wave_base,
                                                                         unsigned* __restrict__ touched) {
    __shared__ unsigned queue[4][128];
    __shared__ float4 lane_rx[4][64];
    if (kStageTop) stage_top_nodes(s);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int segs = (words + 63) >> 6;
    const long long wave = wave_base + (long long)blockIdx.x * 4 + wid;
    if (wave >= (long long)tile_count * segs) return;
    const int tile = (int)(wave / segs);
    const int seg = (int)(wave - (long long)tile * segs);
    const int t = tiles[2 * tile];
    const int base = q.chain_offset[t * kMaxDepth];
    const int end = q.chain_offset[t * kMaxDepth + min(q.max_depth, kMaxDepth)];
    const int tx_words = mask_words(end - base);
    const int w = seg * 64 + lane;
    const unsigned sword = w < tx_words ? super_masks[(size_t)super_of_tile(q, tile) * words + w] : 0u;
    unsigned long long chunks = __ballot(sword != 0u);
    if (chunks == 0ull) return;
    const int first = tiles[2 * tile + 1];
    const int pos = first + lane;
    const bool live = pos < q.cell_offset[t + 1];
    const int cell = live ? cell_at(q, pos) : 0;
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    V3<float> rx{0.0f, 0.0f, 0.0f};
    if (live) rx = ld3(q.rx_pos + 3 * (q.rx_shared ? cell - q.cell_offset[t] : cell));
    const V3<float> lo{wave_min(live ? rx.x : INFINITY), wave_min(live ? rx.y : INFINITY),
                       wave_min(live ? rx.z : INFINITY)};
    const V3<float> hi{wave_max(live ? rx.x : -INFINITY), wave_max(live ? rx.y : -INFINITY),
                       wave_max(live ? rx.z : -INFINITY)};
    lane_rx[wid][lane] = make_float4(rx.x, rx.y, rx.z, __int_as_float(pos));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int queued = 0;  // wave-uniform
    // one round: the first 64 queued pairs are traced, the rest move down
    auto round = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const long long item = lane < queued ? (long long)queue[wid][lane] : -1;
        const long long carry = lane + 64 < queued ? (long long)queue[wid][lane + 64] : -1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (carry >= 0) queue[wid][lane] = (unsigned)carry;
        queued = queued > 64 ? queued - 64 : 0;
        if (item >= 0) {
            const int o = (int)(item & 63), b = (int)((item >> 6) & 63), bb = (int)(item >> 12);
            const float4 r4 = lane_rx[wid][o];
            const V3<float> orx{r4.x, r4.y, r4.z};
            const int chunk = seg * 64 + b;
            const int c = base + chunk * 32 + bb;
            int d = 1;
            while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
            bool ok;
            switch (d) {
                case 1: ok = chain_legs<1, kStageTop>(s, q, c, tx, orx); break;
                case 2: ok = chain_legs<2, kStageTop>(s, q, c, tx, orx); break;
                case 3: ok = chain_legs<3, kStageTop>(s, q, c, tx, orx); break;
                case 4: ok = chain_legs<4, kStageTop>(s, q, c, tx, orx); break;
                default: ok = chain_legs<5, kStageTop>(s, q, c, tx, orx); break;
            }
            if (ok) {
                atomicOr(&masks[(size_t)chunk * q.total_cells + __float_as_int(r4.w)], 1u << bb);
                atomicOr(&touched[(size_t)tile * ((words + 31) >> 5) + (chunk >> 5)], 1u << (chunk & 31));
            }
        }
    };
    int d = 1;
    while (chunks) {
        const int b = __ffsll((long long)chunks) - 1;
        chunks &= chunks - 1;
        const int chunk = seg * 64 + b;
        const unsigned kept = (unsigned)__shfl((int)sword, b);
        const int c0 = base + chunk * 32;
        const int c1 = min(c0 + 32, end);
        bool reach = false;
        if (lane < c1 - c0 && ((kept >> lane) & 1u)) {
            const int c = c0 + lane;
            int dd = 1;
            while (c >= q.chain_offset[t * kMaxDepth + dd]) ++dd;
            reach = chain_reaches_box(s, q, c, dd, lo, hi);
        }
        unsigned todo = (unsigned)__ballot(reach);
        if (todo == 0u) continue;
        // the chunk's words start at zero; the queued legs OR their bits in
        if (live) masks[(size_t)chunk * q.total_cells + pos] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        while (todo) {
            const int bb = __ffs(todo) - 1;
            todo &= todo - 1;
            const int c = c0 + bb;
            while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
            bool lands = false;
            if (live) {
                switch (d) {
                    case 1: lands = chain_lands<1>(s, q, c, tx, rx); break;
                    case 2: lands = chain_lands<2>(s, q, c, tx, rx); break;
                    case 3: lands = chain_lands<3>(s, q, c, tx, rx); break;
                    case 4: lands = chain_lands<4>(s, q, c, tx, rx); break;
                    default: lands = chain_lands<5>(s, q, c, tx, rx); break;
                }
            }
            const unsigned long long ball = __ballot(lands);
            if (lands)
                queue[wid][queued + (int)__popcll(ball & ((1ull << lane) - 1ull))] =
                    (unsigned)(lane | (b << 6) | (bb << 12));
            queued += (int)__popcll(ball);
            if (queued >= 64) round();
        }
    }
    while (queued > 0) round();
}

// Chain validity for per-transmitter point sets (scattered receivers: a
// wavefront's cells gain little from testing the same chain together). After
// the tile cull, each lane back-projects its cell through the surviving
// chains (no traversal); the (cell, chain) pairs that land in faces queue up
// chain by chain (lanes in order within a chain) and their D+1 legs are traced
// 64 pairs per round with every lane busy. The bits are those of
// chain_visibility_kernel (same back-projection and leg code).
template <int D>
__device__ __forceinline__ bool chain_lands(const SceneView& s, const wc_solve& q, int c, const V3<float>& tx,
                                            const V3<float>& rx) {
    V3<double> hd[D];
    int faces[D];
    return solve_chain<D, false>(s, q.chain_groups + c * kMaxDepth, q.chain_images + c * kMaxDepth * 3, tx, rx, hd,
                                 faces);
}

template <int D, bool kTop>
__device__ __forceinline__ bool chain_legs(const SceneView& s, const wc_solve& q, int c, const V3<float>& tx,
                                           const V3<float>& rx) {
    V3<double> hd[D];
    int faces[D];
    const int* groups = q.chain_groups + c * kMaxDepth;
    solve_chain<D, false>(s, groups, q.chain_images + c * kMaxDepth * 3, tx, rx, hd, faces);
    return chain_legs_clear<D, kTop>(s, groups, hd, tx, rx);
}

__global__ void __launch_bounds__(256) chain_visibility_queue_kernel(SceneView s, wc_solve q,
                                                                     const int* __restrict__ tiles, int tile_count,
                                                                     unsigned* __restrict__ masks, int words,
                                                                     const unsigned* __restrict__ super_masks, long long wave_base,
                                                                     unsigned* __restrict__ touched) {
    __shared__ unsigned short queue[4][128];
    __shared__ float4 lane_rx[4][64];
    __shared__ unsigned found[4][64];
    if (kStageTop) stage_top_nodes(s);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long wave = wave_base + (long long)blockIdx.x * 4 + wid;
    const bool wave_live = wave < (long long)tile_count * words;
    const int tile = wave_live ? (int)(wave / words) : 0;
    const int chunk = wave_live ? (int)(wave - (long long)tile * words) : 0;
    const int t = tiles[2 * tile];
    const int width = q.lanes_per_cell;
    const int pos = tiles[2 * tile + 1] + lane;
    const bool live = wave_live && lane < 64 / width && pos < q.cell_offset[t + 1];
    const int cell = live ? cell_at(q, pos) : 0;
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    V3<float> rx{0.0f, 0.0f, 0.0f};
    if (live) rx = ld3(q.rx_pos + 3 * (q.rx_shared ? cell - q.cell_offset[t] : cell));
    const V3<float> lo{wave_min(live ? rx.x : INFINITY), wave_min(live ? rx.y : INFINITY),
                       wave_min(live ? rx.z : INFINITY)};
    const V3<float> hi{wave_max(live ? rx.x : -INFINITY), wave_max(live ? rx.y : -INFINITY),
                       wave_max(live ? rx.z : -INFINITY)};
    const int base = q.chain_offset[t * kMaxDepth];
    const int end = q.chain_offset[t * kMaxDepth + min(q.max_depth, kMaxDepth)];
    const int c0 = base + chunk * 32;
    const int c1 = min(c0 + 32, end);
    // 1. tile cull, one chain per lane, of the chains the tile's supertile
    // kept (a wave whose supertile kept none runs through with no work: the
    // block barriers below need every wave)
    const unsigned kept = (super_masks && wave_live) ? super_masks[(size_t)super_of_tile(q, tile) * words + chunk]
                                                     : ~0u;
    bool reach = false;
    if (wave_live && lane < c1 - c0 && ((kept >> lane) & 1u)) {
        const int c = c0 + lane;
        int d = 1;
        while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
        reach = chain_reaches_box(s, q, c, d, lo, hi);
    }
    unsigned todo = (unsigned)__ballot(reach);
    // 2. back-projection of the lane's cell through the surviving chains
    unsigned cand = 0;
    if (live) {
        unsigned rest = todo;
        int d = 1;
        while (rest) {
            const int b = __ffs(rest) - 1;
            rest &= rest - 1;
            const int c = c0 + b;
            while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
            bool ok;
            switch (d) {
                case 1: ok = chain_lands<1>(s, q, c, tx, rx); break;
                case 2: ok = chain_lands<2>(s, q, c, tx, rx); break;
                case 3: ok = chain_lands<3>(s, q, c, tx, rx); break;
                case 4: ok = chain_lands<4>(s, q, c, tx, rx); break;
                default: ok = chain_lands<5>(s, q, c, tx, rx); break;
            }
            if (ok) cand |= 1u << b;
        }
    }
    // 3. the landing pairs queue up chain by chain; legs traced 64 per round
    lane_rx[wid][lane] = make_float4(rx.x, rx.y, rx.z, 0.0f);
    found[wid][lane] = 0;
    __syncthreads();
    int queued = 0;  // wave-uniform
    const int nb = wave_live ? c1 - c0 : 0;
    for (int b = 0; b <= nb; ++b) {
        if (b < nb) {
            const bool pred = (cand >> b) & 1u;
            const unsigned long long ball = __ballot(pred);
            if (pred) queue[wid][queued + (int)__popcll(ball & ((1ull << lane) - 1ull))] = (unsigned short)(lane | (b << 6));
            queued += (int)__popcll(ball);
            if (queued < 64) continue;
        } else if (queued == 0) {
            break;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int item = lane < queued ? (int)queue[wid][lane] : -1;
        const int carry = lane + 64 < queued ? (int)queue[wid][lane + 64] : -1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (carry >= 0) queue[wid][lane] = (unsigned short)carry;
        queued = queued > 64 ? queued - 64 : 0;
        if (item >= 0) {
            const int o = item & 63, bit = item >> 6;
            const float4 r4 = lane_rx[wid][o];
            const V3<float> orx{r4.x, r4.y, r4.z};
            const int c = c0 + bit;
            int d = 1;
            while (c >= q.chain_offset[t * kMaxDepth + d]) ++d;
            bool ok;
            switch (d) {
                case 1: ok = chain_legs<1, kStageTop>(s, q, c, tx, orx); break;
                case 2: ok = chain_legs<2, kStageTop>(s, q, c, tx, orx); break;
                case 3: ok = chain_legs<3, kStageTop>(s, q, c, tx, orx); break;
                case 4: ok = chain_legs<4, kStageTop>(s, q, c, tx, orx); break;
                default: ok = chain_legs<5, kStageTop>(s, q, c, tx, orx); break;
            }
            if (ok) atomicOr(&found[wid][o], 1u << bit);
        }
    }
    __syncthreads();
    const unsigned bits = live ? found[wid][lane] : 0u;
    if (live) masks[(size_t)chunk * q.total_cells + pos] = bits;
    if (__ballot(bits != 0u) != 0ull && lane == 0 && wave_live)
        atomicOr(&touched[(size_t)tile * ((words + 31) >> 5) + (chunk >> 5)], 1u << (chunk & 31));
}

// Visibility of every (cell, edge) diffraction candidate: bit (ei & 31) of
// word ei >> 5 of the cell's mask is set when the stationary point, clamped
// to the finite edge, sees both the receiver and the transmitter. One
// wavefront covers 64 consecutive evaluation positions x 32 edges (the edge
// loads are wave-uniform). Nothing but geometry and traversal lives here, so
// the kernel keeps few registers and many wavefronts per SIMD in flight to
// hide the traversal's memory latency; the field pass then evaluates the
// visible edges only.
#ifndef WC_VIS_WAVES_PER_EU
#define WC_VIS_WAVES_PER_EU 4
#endif

// Transmitter-leg shadow of one edge cast by one triangle: the edge
// parameter range [0, length] is cut into 64 bins; bit b is set when the
// triangle blocks every segment from the transmitter T to a point of bin b.
// All those segments lie in the plane through T and the edge line; the
// triangle meets that plane in a segment whose central projection from T onto
// the edge line is the interval of blocked edge parameters. The decision
// keeps margins that put every counted crossing well inside the triangle
// (shrunk by 1 mm + 1e-5 of the coordinates), at least 0.1 % of the leg away
// from both of its ends, and at a crossing angle of at least ~6 degrees in the
// plane and between the planes, so the exact segment test (watertight, end
// bias 1e-4 m) blocks those legs too: a set bit only ever removes candidates
// the full test rejects.
struct ShadowFrame {
    V3<double> A, d, w, m;  // edge origin and direction, unit towards T in the fan plane, plane normal
    double a, rho, L, bw, shrink, dt;
    bool valid;
};

__host__ __device__ inline ShadowFrame shadow_frame(const Edge& e, const V3<float>& tx) {
    ShadowFrame f;
    const V3<double> T{tx.x, tx.y, tx.z};
    f.A = {e.ax, e.ay, e.az};
    f.d = {e.dx, e.dy, e.dz};
    f.L = e.length;
    f.a = dot(T - f.A, f.d);
    const V3<double> w = T - f.A - f.d * f.a;
    f.rho = sqrt(dot(w, w));
    f.valid = f.rho > 1.0e-2 && f.L > 0.0;
    f.w = f.valid ? w * (1.0 / f.rho) : V3<double>{0.0, 0.0, 0.0};
    f.m = cross(f.d, f.w);
    const double scale = fabs(f.A.x) + fabs(f.A.y) + fabs(f.A.z) + fabs(T.x) + fabs(T.y) + fabs(T.z) + f.L;
    f.shrink = 1.0e-3 + 1.0e-5 * scale;
    f.bw = f.L / 64.0;
    f.dt = 1.0e-4 + 1.0e-6 * scale;
    return f;
}

__host__ __device__ inline unsigned long long triangle_shadow_bins(const ShadowFrame& f, const Tri& tri) {
    const V3<double> v[3] = {{tri.v0.x, tri.v0.y, tri.v0.z}, {tri.v1.x, tri.v1.y, tri.v1.z},
                             {tri.v2.x, tri.v2.y, tri.v2.z}};
    V3<double> nt = cross(v[1] - v[0], v[2] - v[0]);
    const double nn = sqrt(dot(nt, nt));
    if (!(nn > 0.0)) return 0;
    nt = nt * (1.0 / nn);
    // dihedral angle between the triangle and the fan plane
    const V3<double> c = cross(nt, f.m);
    if (dot(c, c) < 0.01) return 0;
    double sd[3];
    for (int k = 0; k < 3; ++k) sd[k] = dot(f.m, v[k] - f.A);
    V3<double> P[2];
    int np = 0;
    for (int k = 0; k < 3; ++k) {
        const int j = (k + 1) % 3;
        if ((sd[k] < 0.0 && sd[j] > 0.0) || (sd[k] > 0.0 && sd[j] < 0.0)) {
            if (np < 2) P[np] = v[k] + (v[j] - v[k]) * (sd[k] / (sd[k] - sd[j]));
            ++np;
        }
    }
    if (np != 2) return 0;
    // 2-D coordinates in the fan plane: x along the edge, y towards the transmitter
    double x0 = dot(P[0] - f.A, f.d), y0 = dot(P[0] - f.A, f.w);
    double x1 = dot(P[1] - f.A, f.d), y1 = dot(P[1] - f.A, f.w);
    const double len = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
    if (len <= 2.0 * f.shrink) return 0;
    {
        const double ux = (x1 - x0) / len, uy = (y1 - y0) / len;
        x0 += ux * f.shrink;
        y0 += uy * f.shrink;
        x1 -= ux * f.shrink;
        y1 -= uy * f.shrink;
    }
    // clip to the band of crossings at least 0.1 % of the leg from both ends:
    // lambda = 1 - y / rho in [1e-3, 1 - 1e-3]
    const double ylo = 1.0e-3 * f.rho, yhi = (1.0 - 1.0e-3) * f.rho;
    if ((y0 < ylo && y1 < ylo) || (y0 > yhi && y1 > yhi)) return 0;
    if (y0 != y1) {
        const double dy = y1 - y0;
        const double sa = (ylo - y0) / dy, sb = (yhi - y0) / dy;
        const double s0 = fmax(0.0, fmin(sa, sb)), s1 = fmin(1.0, fmax(sa, sb));
        if (!(s0 < s1)) return 0;
        const double cx0 = x0 + (x1 - x0) * s0, cy0 = y0 + dy * s0;
        const double cx1 = x0 + (x1 - x0) * s1, cy1 = y0 + dy * s1;
        x0 = cx0; y0 = cy0; x1 = cx1; y1 = cy1;
    }
    // in-plane crossing angle with the legs at both ends
    const double ex = x1 - x0, ey = y1 - y0, el = sqrt(ex * ex + ey * ey);
    if (!(el > 0.0)) return 0;
    const double ends[2][2] = {{x0, y0}, {x1, y1}};
    for (int k = 0; k < 2; ++k) {
        const double rx = ends[k][0] - f.a, ry = ends[k][1] - f.rho;
        const double rl = sqrt(rx * rx + ry * ry);
        if (!(rl > 0.0) || fabs(ex * ry - ey * rx) < 0.1 * el * rl) return 0;
    }
    // central projection from the transmitter onto the edge line
    const double t0 = f.a + (x0 - f.a) * f.rho / (f.rho - y0);
    const double t1 = f.a + (x1 - f.a) * f.rho / (f.rho - y1);
    const double lo = fmin(t0, t1) + f.dt, hi = fmax(t0, t1) - f.dt;
    if (!(hi > lo)) return 0;
    const double b0 = ceil(lo / f.bw), b1 = floor(hi / f.bw) - 1.0;
    const int first = (int)fmax(b0, 0.0);
    const int last = (int)fmin(b1, 63.0);
    if (first > last) return 0;
    const int count = last - first + 1;
    return (count >= 64 ? ~0ull : ((1ull << count) - 1ull)) << first;
}

// Float pre-test of triangle_shadow_bins. Every leg from T to a point of the
// edge lies in the fan plane, inside the triangle (T, A, A + L d); a blocking
// triangle must cut the plane inside it. The section of a triangle with the
// plane lies within the convex hull of its vertices' in-plane coordinates
// (x along the edge, y towards T), so the test is true when the three
// vertices lie strictly on one side of the plane, or their (x, y) lie
// strictly outside one side of (T, A, B), by more than the frame's shrink
// margin (>= 1 mm). Float distances of scene coordinates below 10 km differ
// from the double ones by well under 1 mm, so the double test would find no
// crossing inside (T, A, B) and return 0 as well.
struct FanPlane {
    float ax, ay, az;  // edge origin
    float mx, my, mz;  // plane normal
    float dx, dy, dz;  // along the edge
    float wx, wy, wz;  // in-plane, towards T
    float a, rho, len, tol;
    float at_x, at_y;  // unit outward normal of side A-T in (x, y)
    float bt_x, bt_y;  // unit outward normal of side B-T
};

__host__ __device__ inline FanPlane fan_plane(const ShadowFrame& f) {
    FanPlane p;
    p.ax = (float)f.A.x; p.ay = (float)f.A.y; p.az = (float)f.A.z;
    p.mx = (float)f.m.x; p.my = (float)f.m.y; p.mz = (float)f.m.z;
    p.dx = (float)f.d.x; p.dy = (float)f.d.y; p.dz = (float)f.d.z;
    p.wx = (float)f.w.x; p.wy = (float)f.w.y; p.wz = (float)f.w.z;
    p.a = (float)f.a; p.rho = (float)f.rho; p.len = (float)f.L; p.tol = (float)f.shrink;
    // side A(0,0)-T(a,rho): outward (away from B(L,0)) normal; side B-T likewise
    const double la = sqrt(f.a * f.a + f.rho * f.rho), lb = sqrt((f.a - f.L) * (f.a - f.L) + f.rho * f.rho);
    p.at_x = (float)(-f.rho / la); p.at_y = (float)(f.a / la);
    p.bt_x = (float)(f.rho / lb); p.bt_y = (float)((f.L - f.a) / lb);
    return p;
}

__host__ __device__ inline bool fan_plane_misses(const FanPlane& p, const Tri& tri) {
    const float vx[3] = {tri.v0.x - p.ax, tri.v1.x - p.ax, tri.v2.x - p.ax};
    const float vy[3] = {tri.v0.y - p.ay, tri.v1.y - p.ay, tri.v2.y - p.ay};
    const float vz[3] = {tri.v0.z - p.az, tri.v1.z - p.az, tri.v2.z - p.az};
    bool above = true, below = true, behind = true, out_a = true, out_b = true;
    for (int k = 0; k < 3; ++k) {
        const float sd = p.mx * vx[k] + p.my * vy[k] + p.mz * vz[k];
        const float x = p.dx * vx[k] + p.dy * vy[k] + p.dz * vz[k];
        const float y = p.wx * vx[k] + p.wy * vy[k] + p.wz * vz[k];
        above = above && sd > p.tol;
        below = below && sd < -p.tol;
        behind = behind && y < -p.tol;
        out_a = out_a && p.at_x * x + p.at_y * y > p.tol;
        out_b = out_b && p.bt_x * (x - p.len) + p.bt_y * y > p.tol;
    }
    return above || below || behind || out_a || out_b;
}

// Shadow bins of every (transmitter, edge) over all triangles outside the
// edge's own face groups (which the leg test ignores). A wavefront serves
// 64 / P edges of one transmitter with P lanes each: lane P j + p takes edge
// j of the group and triangles p, p + P, ..., so a triangle fetch serves
// 64 / P edges (P = 8 when there are many pairs, 64 when few, to fill the
// chip). Only the triangles that reach an edge's fan plane take the
// double-precision section test; the P lanes of an edge then OR their bins.
__global__ void __launch_bounds__(256) edge_shadow_kernel(SceneView s, wc_solve q, int lanes_per_edge,
                                                          unsigned long long* __restrict__ shadow, long long wave_base) {
    const long long wave = wave_base + (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int edges = s.edge_count;
    const int per_wave = 64 / lanes_per_edge;
    const int groups = (edges + per_wave - 1) / per_wave;
    if (wave >= (long long)q.tx_count * groups) return;
    const int lane = threadIdx.x & 63;
    const int t = (int)(wave / groups);
    const int ei = (int)(wave - (long long)t * groups) * per_wave + lane / lanes_per_edge;
    const int phase = lane & (lanes_per_edge - 1);
    unsigned long long bits = 0;
    if (ei < edges) {
        const Edge e = s.edges[ei];
        const ShadowFrame f = shadow_frame(e, ld3(q.tx_pos + 3 * t));
        if (f.valid) {
            const FanPlane plane = fan_plane(f);
            for (int k = phase; k < s.face_count; k += lanes_per_edge) {
                const Tri tri = s.faces[k];
                const int group = __float_as_int(tri.v0.w);
                if (group == e.group0 || group == e.group1 || fan_plane_misses(plane, tri)) continue;
                bits |= triangle_shadow_bins(f, tri);
            }
        }
    }
    for (int off = 1; off < lanes_per_edge; off <<= 1) {
        const unsigned lo32 = __shfl_xor((unsigned)bits, off), hi32 = __shfl_xor((unsigned)(bits >> 32), off);
        bits |= ((unsigned long long)hi32 << 32) | lo32;
    }
    if (ei < edges && phase == 0) shadow[(size_t)t * edges + ei] = bits;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WC_VIS_WAVES_PER_EU)))
diff_visibility_kernel(SceneView s, wc_solve q, unsigned* __restrict__ masks, int words, long long waves,
                       const unsigned long long* __restrict__ shadow, long long wave_base) {
    if (kStageTop) stage_top_nodes(s);
    const long long wave = wave_base + (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= waves) return;
    const int chunk = (int)(wave % words);
    const long long pos64 = (wave / words) * 64 + (threadIdx.x & 63);
    if (pos64 >= q.total_cells) return;
    const int pos = (int)pos64;
    const int t = tx_of_cell(q, pos);
    const int cell = cell_at(q, pos);
    const V3<float> tx = ld3(q.tx_pos + 3 * t);
    const int rx_index = q.rx_shared ? cell - q.cell_offset[t] : cell;
    const V3<float> rx = ld3(q.rx_pos + 3 * rx_index);
    const int e0 = chunk * 32;
    const int e1 = min(e0 + 32, s.edge_count);
    unsigned bits = 0;
    for (int ei = e0; ei < e1; ++ei) {
        const Edge e = s.edges[ei];
        const unsigned long long sh = shadow ? shadow[(size_t)t * s.edge_count + ei] : 0ull;
        if (sh == ~0ull) continue;  // no point of the edge sees the transmitter
        DiffGeometry g;
        if (!diffraction_geometry(e, tx, rx, g, false)) continue;
        if (sh) {
            // the transmitter leg of a stationary point in a shadowed bin is blocked
            const double tc = fmin(fmax(g.tstar, 0.0), (double)e.length);
            const int bin = min((int)(tc * (64.0 / (double)e.length)), 63);
            if ((sh >> bin) & 1ull) continue;
        }
        if (edge_visible<kStageTop>(s, e, g, tx, rx)) bits |= 1u << (ei - e0);
    }
    masks[(size_t)cell * words + chunk] = bits;
}

// The same visibility bits for scattered receivers (per-transmitter point
// sets): a wavefront's receivers are far apart, so lanes gain little from
// tracing the same edge together. Each lane first applies the transmitter-side
// culls to its 32 candidates; the survivors then queue up edge by edge (lanes
// in order within an edge) and are traced 64 per round with every lane busy.
// Each visible bit lands in its owner lane's word through LDS; the bits are
// those of diff_visibility_kernel.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WC_VIS_WAVES_PER_EU)))
diff_visibility_queue_kernel(SceneView s, wc_solve q, unsigned* __restrict__ masks, int words, long long waves,
                       const unsigned long long* __restrict__ shadow, long long wave_base) {
    if (kStageTop) stage_top_nodes(s);
    // per wavefront: the candidate queue (lane | edge << 6), each lane's
    // receiver and transmitter, and the visible bits
    __shared__ unsigned short queue[4][128];
    __shared__ float4 lane_rx[4][64];
    __shared__ unsigned found[4][64];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long wave = wave_base + (long long)blockIdx.x * 4 + wid;
    // waves past the end take part in the barriers with no work
    const bool wave_live = wave < waves;
    const int chunk = wave_live ? (int)(wave % words) : 0;
    const long long pos64 = (wave / words) * 64 + lane;
    const bool live = wave_live && pos64 < q.total_cells;
    const int e0 = chunk * 32;
    const int e1 = min(e0 + 32, s.edge_count);
    int t = 0, cell = 0;
    V3<float> tx{0.0f, 0.0f, 0.0f}, rx{0.0f, 0.0f, 0.0f};
    // 1. the transmitter-side culls of the lane's 32 candidates (no traversal)
    unsigned cand = 0;
    if (live) {
        const int pos = (int)pos64;
        t = tx_of_cell(q, pos);
        cell = cell_at(q, pos);
        tx = ld3(q.tx_pos + 3 * t);
        rx = ld3(q.rx_pos + 3 * (q.rx_shared ? cell - q.cell_offset[t] : cell));
        for (int ei = e0; ei < e1; ++ei) {
            const unsigned long long sh = shadow ? shadow[(size_t)t * s.edge_count + ei] : 0ull;
            if (sh == ~0ull) continue;  // no point of the edge sees the transmitter
            const Edge e = s.edges[ei];
            DiffGeometry g;
            if (!diffraction_geometry(e, tx, rx, g, false)) continue;
            if (sh) {
                // the transmitter leg of a stationary point in a shadowed bin is blocked
                const double tc = fmin(fmax(g.tstar, 0.0), (double)e.length);
                const int bin = min((int)(tc * (64.0 / (double)e.length)), 63);
                if ((sh >> bin) & 1ull) continue;
            }
            cand |= 1u << (ei - e0);
        }
    }
    // 2. candidates queue up edge by edge (lanes in order within an edge),
    // so a round of 64 traces neighbouring cells against the same few edges;
    // a round runs whenever 64 are queued, the remainder at the end
    lane_rx[wid][lane] = make_float4(rx.x, rx.y, rx.z, __int_as_float(t));
    found[wid][lane] = 0;
    __syncthreads();
    int queued = 0;  // wave-uniform
    for (int b = 0; b <= e1 - e0; ++b) {
        if (b < e1 - e0) {
            const bool pred = (cand >> b) & 1u;
            const unsigned long long ball = __ballot(pred);
            if (pred) queue[wid][queued + (int)__popcll(ball & ((1ull << lane) - 1ull))] = (unsigned short)(lane | (b << 6));
            queued += (int)__popcll(ball);
            if (queued < 64) continue;
        } else if (queued == 0) {
            break;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int item = lane < queued ? (int)queue[wid][lane] : -1;
        const int carry = lane + 64 < queued ? (int)queue[wid][lane + 64] : -1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (carry >= 0) queue[wid][lane] = (unsigned short)carry;
        queued = queued > 64 ? queued - 64 : 0;
        if (item >= 0) {
            const int o = item & 63, bit = item >> 6;
            const float4 r4 = lane_rx[wid][o];
            const int ot = __float_as_int(r4.w);
            const V3<float> otx = ld3(q.tx_pos + 3 * ot);
            const V3<float> orx{r4.x, r4.y, r4.z};
            const Edge e = s.edges[e0 + bit];
            DiffGeometry g;
            diffraction_geometry(e, otx, orx, g, false);
            if (edge_visible<kStageTop>(s, e, g, otx, orx)) atomicOr(&found[wid][o], 1u << bit);
        }
    }
    __syncthreads();
    if (live) masks[(size_t)cell * words + chunk] = found[wid][lane];
}

__global__ void __launch_bounds__(256) count_kernel(SceneView s, wc_solve q, int* __restrict__ out) {
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x; cell < q.total_cells; cell += blockDim.x * gridDim.x) {
        const int t = tx_of_cell(q, cell);
        CountSink c;
        eval_cell<CountSink, kPassAll>(s, q, t, cell, 0, 1, c);
        out[cell] = c.count;
    }
}

__global__ void __launch_bounds__(256) emit_kernel(SceneView s, wc_solve q, const int* __restrict__ row_offset,
                                                   const int* __restrict__ path_count, wc_path_rows rows) {
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x; cell < q.total_cells; cell += blockDim.x * gridDim.x) {
        const int t = tx_of_cell(q, cell);
        EmitSink e;
        e.rows = &rows;
        e.row = row_offset[cell];
        e.row_end = e.row + path_count[cell];
        e.tx = t;
        e.rx = cell - q.cell_offset[t];
        eval_cell<EmitSink, kPassAll>(s, q, t, cell, 0, 1, e);
    }
}

// ------------------------------------------------------------------ chain cones
__global__ void chain_cone_kernel(SceneView s, const int* __restrict__ hull_offset,
                                  const double* __restrict__ hull_points, const int* __restrict__ groups,
                                  const double* __restrict__ images, const int* __restrict__ depth, int count,
                                  float* __restrict__ cones) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < count; c += blockDim.x * gridDim.x)
        chain_cone(s.group_plane, hull_offset, hull_points, groups + (size_t)kMaxDepth * c,
                   images + (size_t)kMaxDepth * 3 * c, depth[c], cones + (size_t)kConeFloats * c);
}

// ------------------------------------------------------------------ discovery
__global__ void discover_kernel(SceneView s, const float* __restrict__ tx_pos, int tx_count, int ray_count,
                                int max_depth, int* __restrict__ chains) {
    const long long total = (long long)tx_count * ray_count;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)blockDim.x * gridDim.x) {
        const int t = (int)(idx / ray_count);
        const int i = (int)(idx - (long long)t * ray_count);
        // Fibonacci sphere direction i of ray_count
        const double golden = kPi * (3.0 - sqrt(5.0));
        const double z = 1.0 - (2.0 * i + 1.0) / (double)ray_count;
        const double rad = sqrt(fmax(0.0, 1.0 - z * z));
        const double th = golden * (double)i;
        V3<float> d{(float)(rad * cos(th)), (float)(rad * sin(th)), (float)z};
        V3<float> o = ld3(tx_pos + 3 * t);
        int* out = chains + idx * max_depth;
        int ignore = -1;
        int b = 0;
        for (; b < max_depth; ++b) {
            const RayPre r = ray_setup(o.x, o.y, o.z, d.x, d.y, d.z);
            const Hit h = closest_hit(s, r, 1.0e-4f, 1.0e30f, ignore);
            if (h.t < 0.0f) break;
            out[b] = h.group;
            double off;
            const V3<float> n = to_f(group_normal(s, h.group, off));
            o = o + d * h.t;
            d = d - n * (2.0f * dot(d, n));
            ignore = h.group;
        }
        for (; b < max_depth; ++b) out[b] = -1;
    }
}

__global__ void chain_images_kernel(SceneView s, const float* __restrict__ tx_pos, const int* __restrict__ chain_tx,
                                    const int* __restrict__ chain_groups, const int* __restrict__ chain_depth,
                                    int chain_count, double* __restrict__ images) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < chain_count; c += blockDim.x * gridDim.x) {
        const int t = chain_tx[c];
        V3<double> p{(double)tx_pos[3 * t], (double)tx_pos[3 * t + 1], (double)tx_pos[3 * t + 2]};
        const int d = chain_depth[c];
        for (int j = 0; j < kMaxDepth; ++j) {
            if (j < d) {
                double off;
                const V3<double> n = group_normal(s, chain_groups[c * kMaxDepth + j], off);
                p = p - n * (2.0 * (dot(p, n) - off));
            }
            images[(c * kMaxDepth + j) * 3 + 0] = p.x;
            images[(c * kMaxDepth + j) * 3 + 1] = p.y;
            images[(c * kMaxDepth + j) * 3 + 2] = p.z;
        }
    }
}

__global__ void visibility_kernel(SceneView s, int count, const float* __restrict__ src, const float* __restrict__ dst,
                                  unsigned char* __restrict__ visible) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
        const V3<float> a = ld3(src + 3 * i), b = ld3(dst + 3 * i);
        const V3<float> seg = b - a;
        const float len = norm(seg);
        bool vis = true;
        if (len > 2.0f * kSegmentBias) {
            const float eps = kSegmentBias / len;
            const RayPre r = ray_setup(a.x, a.y, a.z, seg.x, seg.y, seg.z);
            vis = !occluded(s, r, eps, 1.0f - eps, -1, -1);
        }
        visible[i] = vis ? 1 : 0;
    }
}

__global__ void zero_int_kernel(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

static int blocks_for(long long n, int threads) {
    long long b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 1048576) b = 1048576;
    return (int)b;
}

static int check_scene(const wc_scene* scene) {
    if (!scene) return wc_fail("scene is null");
    if (scene->tri_count > 0 && (!scene->nodes || !scene->tris)) return wc_fail("scene BVH pointers are null");
    if (scene->group_count > 0 && (!scene->group_plane || !scene->group_face_offset || !scene->group_faces))
        return wc_fail("scene group tables are null");
    if (scene->edge_count > 0 && !scene->edges) return wc_fail("scene edge table is null");
    return 0;
}

static int check_solve(const wc_solve* q) {
    if (!q) return wc_fail("solve is null");
    if (q->tx_count <= 0 && q->total_cells > 0) return wc_fail("solve has cells but no transmitters");
    if (q->max_depth < 0 || q->max_depth > kMaxDepth) return wc_fail("max_depth must be in [0, 5]");
    if (!(q->frequency_hz > 0.0)) return wc_fail("frequency_hz must be positive");
    if (q->total_cells > 0 && (!q->tx_pos || !q->tx_pol || !q->tx_power || !q->rx_pos || !q->rx_pol ||
                               !q->cell_offset))
        return wc_fail("solve endpoint pointers are null");
    if ((q->components & kCompRefl) && q->max_depth > 0 && (!q->chain_offset || !q->chain_groups || !q->chain_images))
        return wc_fail("reflection requested without chain tables");
    if (q->visibility_mode < 0 || q->visibility_mode > 2) return wc_fail("visibility_mode must be 0, 1 or 2");
    if (q->super_count < 0 || (size_t)q->super_count > super_capacity(q->total_cells, q->tx_count))
        return wc_fail("super_count %d exceeds ceil(total_cells / 1024) + tx_count", q->super_count);
    if (q->super_count > 0 && (!q->super_first_tile || !q->tile_super))
        return wc_fail("super_count > 0 needs super_first_tile and tile_super");
    return 0;
}

}  // namespace wc

using namespace wc;

// Zero fill (coalesced 4-byte stores; a fill kernel, not a memset node, keeps
// the launch graph-capturable).
__global__ void __launch_bounds__(256) fill_zero_kernel(unsigned* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Launch-size guard of the wavefront-indexed kernels: a grid covers at most
// kWavesPerLaunch wavefronts (4 per 256-thread block, 2^28 threads), larger
// wave counts run as consecutive launches with their first wave passed as
// wave_base. Every launch's status is kept (the first failure is reported by
// the entry point even when later launches succeed).
constexpr long long kWavesPerLaunch = 1LL << 22;
#define WC_LAUNCH(...)                 \
    do {                               \
        hipLaunchKernelGGL(__VA_ARGS__); \
        note_launch();                 \
    } while (0)
thread_local hipError_t t_launch_error = hipSuccess;

static void note_launch() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess && t_launch_error == hipSuccess) t_launch_error = e;
}

template <class Launch>
static void for_wave_chunks(long long waves, Launch&& launch) {
    for (long long base = 0; base < waves; base += kWavesPerLaunch) {
        const long long n = std::min(kWavesPerLaunch, waves - base);
        launch(dim3((unsigned)((n + 3) / 4)), base);
        note_launch();
    }
}

// chain words (32 chains each) from which the supertile cull pays for its
// two extra launches
constexpr int kSuperMinWords = 64;

// the tile loops (neighbouring receivers per wavefront) or the queued legs
__host__ inline bool coherent_tiles(const wc_solve& q) {
    return q.visibility_mode == 1 || (q.visibility_mode == 0 && q.rx_shared) || q.lanes_per_cell != 1;
}

static void launch_chain_visibility(const SceneView& s, const wc_solve& q, const wc_radiomap_out& o, const int* tiles,
                                    int tile_count, hipStream_t st) {
    if (q.max_tx_chains <= 0 || !(q.components & kCompRefl) || q.max_depth <= 0) return;
    const int words = mask_words(q.max_tx_chains);
    unsigned* masks = chain_masks(o.workspace, q, s.edge_count);
    unsigned* touched = touched_bitmaps(o.workspace, q, s.edge_count);
    const long long waves = (long long)tile_count * words;
    // every tile's chunk bitmap starts empty; the mask words of a chunk are
    // written (or zeroed, then OR-ed) only where the bitmap marks the chunk
    const size_t touched_count = (size_t)tile_count * (size_t)touched_words(q.max_tx_chains);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<size_t>((touched_count + 255) / 256, 65536)),
                       dim3(256), 0, st, touched, touched_count);
    note_launch();
    // many chains: cull them per supertile first (host supertile tables, or
    // for one transmitter the 1,024-position blocks in order); the tile pass
    // then visits only what its supertile kept
    const unsigned* super_masks = nullptr;
    const bool tables = q.super_first_tile && q.tile_super && q.super_count > 0;
    if (q.lanes_per_cell == 1 && words >= kSuperMinWords && (tables || q.tx_count == 1)) {
        const SuperLayout l = super_layout(
            shadow_end(q.total_cells, q.lanes_per_cell, s.edge_count, q.max_tx_chains, q.tx_count), q.total_cells,
            q.max_tx_chains, q.tx_count);
        char* ws = static_cast<char*>(o.workspace);
        float* boxes = reinterpret_cast<float*>(ws + l.boxes);
        unsigned* sm = reinterpret_cast<unsigned*>(ws + l.masks);
        wc_solve qs = q;
        if (!tables) {
            qs.super_first_tile = nullptr;
            qs.tile_super = nullptr;
        }
        const int supers = tables ? q.super_count : (q.total_cells + kSuperCells - 1) / kSuperCells;
        hipLaunchKernelGGL(super_box_kernel, dim3((unsigned)((supers + 3) / 4)), dim3(256), 0, st, qs, tiles, supers,
                           boxes);
        note_launch();
        for_wave_chunks((long long)supers * words, [&](dim3 grid, long long base) {
            hipLaunchKernelGGL(super_cull_kernel, grid, dim3(256), 0, st, s, qs, tiles, supers, words, boxes, sm, base);
        });
        super_masks = sm;
        if (coherent_tiles(q))
            for_wave_chunks((long long)tile_count * ((words + 63) / 64), [&](dim3 grid, long long base) {
                hipLaunchKernelGGL(chain_visibility_seg_kernel, grid, dim3(256), 0, st, s, qs, tiles, tile_count,
                                   masks, words, super_masks, base, touched);
            });
        else
            for_wave_chunks((long long)tile_count * ((words + 63) / 64), [&](dim3 grid, long long base) {
                hipLaunchKernelGGL(chain_visibility_seg_queue_kernel, grid, dim3(256), 0, st, s, qs, tiles,
                                   tile_count, masks, words, super_masks, base, touched);
            });
        return;
    }
    if (coherent_tiles(q))
        for_wave_chunks(waves, [&](dim3 grid, long long base) {
            hipLaunchKernelGGL(chain_visibility_kernel, grid, dim3(256), 0, st, s, q, tiles, tile_count, masks, words,
                               super_masks, base, touched);
        });
    else
        for_wave_chunks(waves, [&](dim3 grid, long long base) {
            hipLaunchKernelGGL(chain_visibility_queue_kernel, grid, dim3(256), 0, st, s, q, tiles, tile_count, masks,
                               words, super_masks, base, touched);
        });
}

// Compute-unit count of the current device, queried once per device and
// thread (safe during stream capture).
static int device_cus() {
    static thread_local int cached_device = -1;
    static thread_local int cached_cus = 256;
    int device = 0;
    (void)hipGetDevice(&device);
    if (device != cached_device) {
        int cus_query = 256;
        if (hipDeviceGetAttribute(&cus_query, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess)
            cached_cus = cus_query;
        cached_device = device;
    }
    return cached_cus;
}

// The shadow table costs transmitters x edges x triangles segment-plane
// tests; above this many it is skipped (the visibility test alone is exact).
constexpr long long kShadowBudget = 4000000000LL;

static void launch_diff_visibility(const SceneView& s, const wc_solve& q, const wc_radiomap_out& o, hipStream_t st) {
    const int words = mask_words(s.edge_count);
    if (words == 0) return;
    unsigned* masks = reinterpret_cast<unsigned*>(static_cast<char*>(o.workspace) +
                                                  partial_bytes(q.total_cells, q.lanes_per_cell));
    unsigned long long* shadow = nullptr;
    const long long pairs = (long long)q.tx_count * s.edge_count;
    if (pairs > 0 && pairs * (long long)s.face_count <= kShadowBudget) {
        shadow = reinterpret_cast<unsigned long long*>(
            static_cast<char*>(o.workspace) + shadow_offset(q.total_cells, q.lanes_per_cell, s.edge_count,
                                                            q.max_tx_chains, q.tx_count));
        // 64 lanes per edge: on the urban canyon with 64 transmitters, 8 lanes
        // per edge (fewer triangle fetches) measured slower
        const int lanes_per_edge = 64;
        const int per_wave = 64 / lanes_per_edge;
        const long long shadow_waves = (long long)q.tx_count * ((s.edge_count + per_wave - 1) / per_wave);
        for_wave_chunks(shadow_waves, [&](dim3 grid, long long base) {
            hipLaunchKernelGGL(edge_shadow_kernel, grid, dim3(256), 0, st, s, q, lanes_per_edge, shadow, base);
        });
    }
    const long long waves = ((long long)q.total_cells + 63) / 64 * words;
    // a receiver grid shared by the transmitters keeps neighbouring receivers
    // in a wavefront (the edge-by-edge loop traces coherently: 0.84 vs 0.97 ms
    // on the 256 x 256 canyon); per-transmitter point sets take the queue
    // (1.16 vs 1.26 ms on 64 transmitters with scattered street points)
    const unsigned long long* sh = shadow;
    if (coherent_tiles(q))
        for_wave_chunks(waves, [&](dim3 grid, long long base) {
            hipLaunchKernelGGL(diff_visibility_kernel, grid, dim3(256), 0, st, s, q, masks, words, waves, sh, base);
        });
    else
        for_wave_chunks(waves, [&](dim3 grid, long long base) {
            hipLaunchKernelGGL(diff_visibility_queue_kernel, grid, dim3(256), 0, st, s, q, masks, words, waves, sh,
                               base);
        });
}

// Launch stages: bit 0 chain validity (tiled only), bit 1 the
// reflection/LoS/transmission field pass, bit 2 edge visibility, bit 3 the
// diffraction field pass.
enum Stage : int { kStageChains = 1, kStageMain = 2, kStageEdges = 4, kStageDiff = 8, kStageAll = 15 };

static int radiomap_launch(const wc_scene* scene, const wc_solve* solve, const wc_radiomap_out* out,
                           const int* tiles, int tile_count, int* tile_counter, int stages, const char* who,
                           void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (int rc = check_solve(solve)) return rc;
    if (!out || !out->power_total || !out->field_total || !out->component_power || !out->component_field)
        return wc_fail("%s: output pointers are null", who);
    if (out->topology && out->topology_capacity <= 0) return wc_fail("topology_capacity must be positive");
    const int width = solve->lanes_per_cell;
    if (width != 1 && width != 2 && width != 4 && width != 8 && width != 16 && width != 32 && width != 64)
        return wc_fail("lanes_per_cell must be a power of two in [1, 64]");
    if (!tiles && width != 1) return wc_fail("lanes_per_cell > 1 needs the tiled launch");
    if (out->topology && out->topology_capacity % width != 0)
        return wc_fail("topology_capacity must be a multiple of lanes_per_cell");
    const bool diffraction = (solve->components & kCompDiff) != 0;
    const bool chains = tiles && solve->max_tx_chains > 0 && (solve->components & kCompRefl) && solve->max_depth > 0;
    if (!diffraction) stages &= ~(kStageEdges | kStageDiff);
    if (!chains) stages &= ~kStageChains;
    if (solve->total_cells == 0 || stages == 0) return 0;
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipStream_t st = (hipStream_t)stream;
    (void)hipGetLastError();
    t_launch_error = hipSuccess;
    if ((diffraction || chains) && !out->workspace)
        return wc_fail("%s: workspace (wc_radiomap_workspace_bytes) is required with diffraction or reflection", who);
    if (tiles) {
        if (!tile_counter || tile_count <= 0) return wc_fail("tiled launch needs tile_counter and tile_count");
        if ((size_t)tile_count > tile_capacity(solve->total_cells, width, solve->tx_count))
            return wc_fail("%s: tile_count %d exceeds ceil(total_cells * lanes_per_cell / 64) + tx_count", who,
                           tile_count);
        const int cached_cus = device_cus();
        int blocks = std::min((tile_count + 3) / 4, cached_cus * 8);
        blocks = std::max(blocks, 1);
        // a fill kernel rather than a memset node keeps the launches graph-capturable
        const bool main = (stages & kStageMain) != 0, diff = (stages & kStageDiff) != 0;
        if (main && diff)
            WC_LAUNCH(zero_int_kernel, dim3(1), dim3(64), 0, st, tile_counter, 2);
        else if (main || diff)
            WC_LAUNCH(zero_int_kernel, dim3(1), dim3(64), 0, st, tile_counter + (main ? 0 : 1), 1);
        if (stages & kStageChains) launch_chain_visibility(s, *solve, *out, tiles, tile_count, st);
        if (main)
            WC_LAUNCH(radiomap_main_tiles, dim3(blocks), dim3(256), 0, st, s, *solve, *out, tiles,
                               tile_count, tile_counter);
        if (stages & kStageEdges) launch_diff_visibility(s, *solve, *out, st);
        if (diff) {
            const int* staged = nullptr;
            const int words = mask_words(s.edge_count);
            const int n = solve->total_cells;
            if (width == 1 && words > 0) {
                const DiffStageLayout l = diff_stage_layout(
                    stage_base(n, width, s.edge_count, solve->max_tx_chains, solve->tx_count), n, s.edge_count);
                char* ws = static_cast<char*>(out->workspace);
                int* offsets = reinterpret_cast<int*>(ws + l.offsets);
                int* ctrl = reinterpret_cast<int*>(ws + l.ctrl);
                DiffRecord* records = reinterpret_cast<DiffRecord*>(ws + l.records);
                const unsigned* masks = reinterpret_cast<const unsigned*>(ws + partial_bytes(n, width));
                WC_LAUNCH(diff_count_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, *solve, masks,
                                   words, offsets);
                WC_LAUNCH(diff_scan_kernel, dim3(1), dim3(1024), 0, st, offsets, n,
                                   diff_record_capacity(n, s.edge_count), ctrl);
                WC_LAUNCH(diff_record_kernel, dim3(cached_cus * 4), dim3(256), 0, st, s, *solve, masks,
                                   words, offsets, ctrl, records);
                WC_LAUNCH(diff_sum_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, *solve, *out,
                                   offsets, ctrl, records);
                staged = ctrl;
            }
            WC_LAUNCH(radiomap_diff_tiles, dim3(blocks), dim3(256), 0, st, s, *solve, *out, tiles,
                               tile_count, tile_counter + 1, staged);
        }
    } else {
        if (stages & kStageMain)
            WC_LAUNCH(radiomap_kernel<kPassMain>, dim3(blocks_for(solve->total_cells, 256)), dim3(256), 0,
                               st, s, *solve, *out);
        if (stages & kStageEdges) launch_diff_visibility(s, *solve, *out, st);
        if (stages & kStageDiff)
            WC_LAUNCH(radiomap_kernel<kPassDiff>, dim3(blocks_for(solve->total_cells, 256)), dim3(256), 0,
                               st, s, *solve, *out);
    }
    note_launch();
    const hipError_t err = t_launch_error;
    t_launch_error = hipSuccess;
    return err == hipSuccess ? 0 : wc_fail_hip(who, err);
}

extern "C" int wc_radiomap_forward(const wc_scene* scene, const wc_solve* solve, const wc_radiomap_out* out,
                                   const int* tiles, int tile_count, int* tile_counter, void* stream) {
    return radiomap_launch(scene, solve, out, tiles, tile_count, tile_counter, kStageAll, "wc_radiomap_forward",
                           stream);
}

extern "C" int wc_radiomap_pass(const wc_scene* scene, const wc_solve* solve, const wc_radiomap_out* out,
                                const int* tiles, int tile_count, int* tile_counter, int pass, void* stream) {
    if (pass < 1 || pass > 4) return wc_fail("wc_radiomap_pass: pass must be 1, 2, 3 or 4 (got %d)", pass);
    if (!tiles) return wc_fail("wc_radiomap_pass: the pass-by-pass launch needs tiles");
    if (pass >= 3 && !(solve && (solve->components & kCompDiff)))
        return wc_fail("wc_radiomap_pass: passes 3 and 4 need a solve with diffraction");
    return radiomap_launch(scene, solve, out, tiles, tile_count, tile_counter, 1 << (pass - 1), "wc_radiomap_pass",
                           stream);
}

extern "C" size_t wc_radiomap_workspace_bytes(int cells, int lanes_per_cell, int edge_count, int max_tx_chains,
                                              int tx_count) {
    const int c = cells > 0 ? cells : 0;
    const int e = edge_count > 0 ? edge_count : 0;
    const int lanes = lanes_per_cell > 0 ? lanes_per_cell : 1;
    const size_t base = stage_base(c, lanes, e, max_tx_chains, tx_count);
    // the staged diffraction pass runs with one lane per cell only
    return lanes == 1 && e > 0 ? diff_stage_layout(base, c, e).end : base;
}

extern "C" int wc_paths_count(const wc_scene* scene, const wc_solve* solve, int* path_count, void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (int rc = check_solve(solve)) return rc;
    if (!path_count && solve->total_cells > 0) return wc_fail("wc_paths_count: path_count is null");
    if (solve->total_cells == 0) return 0;
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(count_kernel, dim3(blocks_for(solve->total_cells, 256)), dim3(256), 0, (hipStream_t)stream, s,
                       *solve, path_count);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_paths_count", err);
}

extern "C" int wc_paths_emit(const wc_scene* scene, const wc_solve* solve, const int* row_offset,
                             const int* path_count, const wc_path_rows* rows, void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (int rc = check_solve(solve)) return rc;
    if (!rows || !row_offset || !path_count) return wc_fail("wc_paths_emit: rows / row_offset / path_count are null");
    if (rows->width < 1 || rows->width < solve->max_depth) return wc_fail("wc_paths_emit: width must be >= max_depth");
    if (solve->total_cells == 0) return 0;
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(emit_kernel, dim3(blocks_for(solve->total_cells, 256)), dim3(256), 0, (hipStream_t)stream, s,
                       *solve, row_offset, path_count, *rows);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_paths_emit", err);
}

extern "C" int wc_chain_cones(const wc_scene* scene, const int* hull_offset, const double* hull_points,
                              const int* groups, const double* images, const int* depth, int count, float* cones,
                              void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (count < 0) return wc_fail("wc_chain_cones: count must be non-negative");
    if (count == 0) return 0;
    if (!hull_offset || !hull_points || !groups || !images || !depth || !cones)
        return wc_fail("wc_chain_cones: null argument");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(chain_cone_kernel, dim3(blocks_for(count, 64)), dim3(64), 0, (hipStream_t)stream, s,
                       hull_offset, hull_points, groups, images, depth, count, cones);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_chain_cones", err);
}

extern "C" int wc_visibility(const wc_scene* scene, int count, const float* source, const float* target,
                             unsigned char* visible, void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (count < 0) return wc_fail("wc_visibility: count must be non-negative");
    if (count == 0) return 0;
    if (!source || !target || !visible) return wc_fail("wc_visibility: null argument");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(visibility_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream, s, count,
                       source, target, visible);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_visibility", err);
}

extern "C" int wc_discover_chains(const wc_scene* scene, const float* tx_pos, int tx_count, int ray_count,
                                  int max_depth, int* chains, void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (tx_count < 0 || ray_count < 0) return wc_fail("wc_discover_chains: counts must be non-negative");
    if (max_depth < 1 || max_depth > kMaxDepth) return wc_fail("wc_discover_chains: max_depth must be in [1, 5]");
    const long long total = (long long)tx_count * ray_count;
    if (total == 0) return 0;
    if (!tx_pos || !chains) return wc_fail("wc_discover_chains: null argument");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(discover_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, (hipStream_t)stream, s, tx_pos,
                       tx_count, ray_count, max_depth, chains);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_discover_chains", err);
}

extern "C" int wc_chain_images(const wc_scene* scene, const float* tx_pos, const int* chain_tx,
                               const int* chain_groups, const int* chain_depth, int chain_count, double* images,
                               void* stream) {
    if (int rc = check_scene(scene)) return rc;
    if (chain_count < 0) return wc_fail("wc_chain_images: chain_count must be non-negative");
    if (chain_count == 0) return 0;
    if (!tx_pos || !chain_tx || !chain_groups || !chain_depth || !images) return wc_fail("wc_chain_images: null argument");
    const SceneView s = *reinterpret_cast<const SceneView*>(scene);
    hipLaunchKernelGGL(chain_images_kernel, dim3(blocks_for(chain_count, 256)), dim3(256), 0, (hipStream_t)stream, s,
                       tx_pos, chain_tx, chain_groups, chain_depth, chain_count, images);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : wc_fail_hip("wc_chain_images", err);
}
