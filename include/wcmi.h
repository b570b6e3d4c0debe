/*
 * wcmi.h - C ABI of libwcmi.so, the MI355X (gfx950) implementation of the
 * witwin-channel deterministic / path propagation hot path.
 *
 * Conventions
 *  - Every pointer argument is a device pointer (HBM) unless the name ends in
 *    `_host`. The caller owns and allocates all buffers (PyTorch tensors on
 *    the Python side); the library never allocates device memory except in
 *    wc_scene_* helpers that are documented to do so.
 *  - `stream` is a hipStream_t passed as void*. Work is enqueued
 *    asynchronously on it; no entry point synchronizes unless stated.
 *  - Return value 0 = success. Non-zero = failure; wc_last_error() returns a
 *    thread-local message. Host-side validation failures return 1, HIP
 *    launch errors 2.
 *  - Vectors are float32 xyz triples, complex numbers are interleaved
 *    (re, im) float32 pairs (the memory layout of torch.complex64).
 *
 * Each entry point names the reference operator it replaces (file:line in
 * the witwin-channel source tree).
 */
#ifndef WCMI_H
#define WCMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ identity */
const char* wc_last_error(void);
void wc_clear_error(void);
int wc_abi_version(void);
const char* wc_build_target(void);

/* ------------------------------------------------------------ scene / LBVH
 * Replaces the RayD scene creation behind _channel.rayd_scene_create
 * (native/channel/binding/rayd.cpp:367, native/channel/rayd/scene.cpp:19-57):
 * builds a linear BVH (Morton codes, LDS radix sort, Karras hierarchy,
 * bottom-up boxes) over the triangles. The output is the BINARY tree
 * (64 B nodes); the scene's nodes are its 4-wide collapse (wc_bvh_widen).
 *   nodes_out:        wc_bvh_node_count(face_count) x 64 B
 *   sorted_tris_out:  face_count x 48 B (BVH leaf order)
 *   face_tris_out:    face_count x 48 B (original face order)
 *   workspace:        wc_bvh_workspace_bytes(face_count) bytes
 */
size_t wc_bvh_workspace_bytes(int face_count);
int wc_bvh_node_count(int face_count);
int wc_bvh_build(const float* vertices, const int* faces, const int* face_group, const int* face_material,
                 int face_count, void* nodes_out, void* sorted_tris_out, void* face_tris_out, void* workspace,
                 size_t workspace_bytes, void* stream);
/* The same scene creation (rayd.cpp:367, scene.cpp:19-57) with a binned
 * surface-area-heuristic tree instead: 16 centroid bins per axis, one
 * triangle per leaf, collapsed to the 4-wide nodes the tracers walk. Host
 * pointers in and out (upload the three tables afterwards).
 *   nodes_out_host:   wc_bvh_node_count(face_count) x 64 B (4-wide nodes;
 *                     slots past the tree's node count stay untouched)
 *   sorted_tris_out_host / face_tris_out_host: as for wc_bvh_build
 * Every visibility and hit query answers as with the LBVH; a walk visits
 * about a third fewer nodes on building and terrain meshes. */
int wc_bvh_build_sah(const float* vertices_host, const int* faces_host, const int* face_group_host,
                     const int* face_material_host, int face_count, void* nodes_out_host,
                     void* sorted_tris_out_host, void* face_tris_out_host);
/* Collapse of a binary tree (wc_bvh_build's nodes, copied to the host) into
 * the 4-wide nodes of the scene view: each wide node takes a binary node's
 * children and opens the internal child of largest surface area until it
 * holds four (half the dependent node fetches per walk), its child boxes as
 * 8-bit steps of a per-axis grid chosen so every decoded bound is an exact
 * float (64 B per node, one cache line). Fails, rather than
 * let a walk overflow, when no collapse fits the 96-entry walk stack. Host
 * pointers; wide_nodes_out_host holds wc_bvh_node_count(face_count) x 64 B;
 * wide_count_host (optional) receives the wide node count. */
int wc_bvh_widen(const void* binary_nodes_host, int face_count, void* wide_nodes_out_host, int* wide_count_host);

/* Scene view consumed by the path kernels. Group planes are float64 (unit
 * normal xyz, offset). Material arrays hold the first-layer scalar view;
 * layer arrays hold the CSR stacks (reference scene/compiler.py:131-261).
 * Edge records are 80-byte structs laid out as in wc_edge. */
typedef struct wc_edge {
    float ax, ay, az;
    float dx, dy, dz;
    float length;
    float n;
    float t0x, t0y, t0z;
    float n0x, n0y, n0z;
    int face0, face1;
    int group0, group1;
    int material0, material1;
} wc_edge;

typedef struct wc_scene {
    const void* nodes;
    const void* tris;
    int tri_count;
    int node_count;
    const void* faces;
    const int* face_group;
    const int* face_material;
    int face_count;
    const double* group_plane;
    const int* group_face_offset;
    const int* group_faces;
    const int* group_material;
    int group_count;
    const float* mat_eps;
    const float* mat_sigma;
    const float* mat_mu;
    const float* mat_gain;
    const float* mat_thickness;
    const int* mat_layer_offset;
    const int* mat_layer_count;
    const float* layer_thickness;
    const float* layer_eps;
    const float* layer_sigma;
    const float* layer_mu;
    int material_count;
    const wc_edge* edges;
    int edge_count;
} wc_scene;

/* Transmitters, receiver cells and candidate reflection chains of one solve.
 * Cells of transmitter t are [cell_offset[t], cell_offset[t+1]); with
 * rx_shared = 1 every transmitter uses rx_pos[cell - cell_offset[t]], else
 * rx_pos[cell]. rx_pol is [1,3] (rx_pol_stride 0) or one per position.
 * Chains of (transmitter t, depth d) are rows chain_offset[t*5+d-1] ..
 * chain_offset[t*5+d] of chain_groups [K,5] / chain_images [K,5,3] (float64). */
typedef struct wc_solve {
    const float* tx_pos;
    const float* tx_pol;
    const float* tx_power;
    int tx_count;
    const float* rx_pos;
    const float* rx_pol;
    int rx_pol_stride;
    const int* cell_offset;
    int rx_shared;
    int total_cells;
    const int* chain_offset;
    const int* chain_groups;
    const double* chain_images;
    double frequency_hz; /* carrier (Hz); the phase -k L uses it in double */
    int components;  /* bit 0 los, 1 reflection, 2 diffraction, 3 transmission; bit 4: a
                        transmitter-receiver segment crossing more than max_depth walls sets
                        failure bit 4 (wc_radiomap_out.failure) instead of carrying no path */
    int max_depth;
    int coherent;
    int lanes_per_cell; /* L: lanes sharing one cell's candidates (power of two, 1..64);
                           tiles then hold 64 / L cells */
    const int* cell_order; /* [total_cells] or NULL (identity): evaluation position p
                              works on cell cell_order[p]; a permutation that maps each
                              transmitter's range [cell_offset[t], cell_offset[t+1]) onto
                              itself. Tiles and launches run over positions, outputs stay
                              indexed by cell; a spatially coherent order (receivers
                              that are close together in one wavefront) cuts divergence. */
    int max_tx_chains;     /* largest reflection-chain count of one transmitter (all
                              depths); sizes the per-cell chain masks of the tiled
                              radiomap (0: no chain masks) */
    /* Supertiles of the chain validity cull (optional; NULL / 0: one
     * transmitter's supertiles are its 16 consecutive tiles by position):
     * super s covers tiles [super_first_tile[s], super_first_tile[s+1]) of the
     * tile list, at most 16 tiles of one transmitter; tile_super[tile] is the
     * super holding a tile. */
    const int* super_first_tile; /* [super_count + 1] */
    const int* tile_super;       /* [tile_count] */
    int super_count;             /* <= ceil(total_cells / 1024) + tx_count */
    int visibility_mode;  /* 0: by receiver layout (shared grid -> tile loops,
                             per-transmitter sets -> queued legs); 1: tile loops
                             (neighbouring receivers per wavefront); 2: queued legs */
    const float* chain_cones; /* [K, WC_CONE_FLOATS] receiver cones of the chains
                                 (wc_chain_cones) or NULL: the chain culls then test
                                 each chain's cone before its last-bounce faces */
} wc_solve;

/* ------------------------------------------------------------ scene refit
 * Moved vertices, same topology: replaces the per-world-version recompile and
 * RayD scene rebuild (scene/compiler.py:1306-1450, scene/resources.py:169-245)
 * for geometry steps such as a vertex-gradient loop. In place, on `stream`:
 * the triangle tables (tris, faces) get the new positions, every 4-wide node
 * is refit after its children (levels deepest first) and its child boxes are
 * re-quantized on the exact grid of the host build, group planes are
 * recomputed from their representative faces and edge records from their four
 * vertices (float64, the host compile's formulas). The tree's shape, groups
 * and edge list stay; walks stay exact, they may visit more nodes as the
 * boxes drift from a fresh build. A box that cannot be quantized sets
 * *failure (the tree is then unusable: rebuild).
 *   node_order:        wide node ids grouped by depth, deepest level first
 *   level_offset_host: [levels + 1] host offsets into node_order */
typedef struct wc_refit {
    const float* vertices;        /* [V,3] new positions */
    const int* faces;             /* [F,3] */
    const int* group_face;        /* [G] representative face of each group */
    const int* edge_vertices;     /* [E,4] (a, b, c0, c1), c1 = -1 for a half-plane */
    const int* node_order;
    const int* level_offset_host;
    int levels;
    int* failure;                 /* device int, caller-zeroed */
} wc_refit;
int wc_scene_refit(const wc_scene* scene, const wc_refit* refit, void* stream);

/* ------------------------------------------------------------ radiomap
 * Fused deterministic radiomap: per (transmitter, cell) discovery of every
 * line-of-sight, reflection chain, diffraction edge and transmission path,
 * field evaluation, and coherent accumulation in float64 registers. No path
 * rows are materialized. Replaces the chain evaluate_enumerated_paths ->
 * enumerated_canonical_compact -> evaluate_path_fields ->
 * deterministic_accumulate_flat
 * (witwin/channel/propagation/enumerated.py:549-837,
 *  native/channel/kernels/deterministic.cu:860-970).
 * Outputs (cells = total_cells):
 *   power_total [cells] f32, field_total [cells] c64,
 *   component_power [4, cells] f32 (slots los, reflection, diffraction,
 *   transmission), component_field [4, cells] c64,
 *   diffraction_vector [cells, 3] c64 (may be NULL),
 *   path_count [cells] i32 (may be NULL),
 *   topology [cells, topology_capacity] i32 path descriptors for the
 *   fixed-topology adjoint (may be NULL); failure [1] i32 gets bit 1 when a
 *   cell holds more paths than topology_capacity.
 * tiles != NULL: persistent launches over wave-aligned tiles (tx, first cell)
 * of 64 cells pulled from global work counters (tile_counter [2] i32, zeroed
 * by the call with a fill kernel, so the sequence is graph-capturable).
 * The evaluation runs as launches with their own register budgets (tiled
 * form): a chain-validity launch that back-projects every (cell, reflection
 * chain) candidate and traces its legs into bit masks, the
 * reflection/LoS/transmission field launch over the valid chains, then (with
 * diffraction) a visibility launch that tests both legs of every (cell, edge)
 * candidate into bit masks and the diffraction field launch over the visible
 * edges. */
typedef struct wc_radiomap_out {
    float* power_total;
    float* field_total;
    float* component_power;
    float* component_field;
    float* diffraction_vector;
    int* path_count;
    int* topology;
    int topology_capacity;
    int* failure;
    void* workspace; /* wc_radiomap_workspace_bytes(cells, L, edge_count, max_tx_chains,
                        tx_count): per-lane sums carried from the reflection pass to
                        the diffraction pass, one edge-visibility bit per (cell, edge),
                        one chain-validity bit per (cell, chain) and 64 transmitter-leg
                        shadow bins per (transmitter, edge) */
} wc_radiomap_out;

size_t wc_radiomap_workspace_bytes(int cells, int lanes_per_cell, int edge_count, int max_tx_chains,
                                   int tx_count);
int wc_radiomap_forward(const wc_scene* scene, const wc_solve* solve, const wc_radiomap_out* out,
                        const int* tiles, int tile_count, int* tile_counter, void* stream);

/* One launch of wc_radiomap_forward's tiled sequence, for callers that time
 * or schedule its kernels apart: pass 1 = chain validity (reflection chain
 * masks), 2 = reflection/LoS/transmission field pass (zeroes
 * tile_counter[0]; reads the chain masks), 3 = edge visibility (diffraction
 * masks), 4 = diffraction field pass (zeroes tile_counter[1]; reads the
 * per-cell sums pass 2 left in `out->workspace` and the edge masks). Run in
 * that order on one solve and output set, passes 1-4 (3 and 4 only with
 * diffraction) equal wc_radiomap_forward. No reference counterpart: the
 * reference runs its radiomap as one sequence of launches. */
int wc_radiomap_pass(const wc_scene* scene, const wc_solve* solve, const wc_radiomap_out* out,
                     const int* tiles, int tile_count, int* tile_counter, int pass, void* stream);

/* ------------------------------------------------------------ radiomap adjoint
 * Fixed-topology VJP of wc_radiomap_forward. Replaces the reference adjoint
 * chain _DeterministicAccumulateFlatAdFunction.backward ->
 * _Field*AdFunction.backward -> RayD EPC backward
 * (witwin/channel/kernels/deterministic.py:186-367, kernels/fields.py:1087-1827,
 * native/channel/kernels/deterministic.cu:977-1117, fields.cu:1237,1873).
 * Instead of tapes it recomputes every path stored in `topology` (from the
 * forward call) with forward-mode dual numbers over the path's transmitter
 * position, frequency and per-interaction material parameters, contracts
 * with the cell adjoint, and writes one partial gradient vector per work
 * item (a slice of one tile's paths):
 * partials [wc_vjp_partial_rows(tile_count), wc_vjp_partial_stride(M)] =
 *   [dL/deps_r[0..M), dL/dsigma[0..M), dL/df, dL/dtx_x, dL/dtx_y, dL/dtx_z]
 * where the transmitter terms belong to the item's tile's transmitter (row r
 * belongs to tile r / (rows / tile_count)). The upstream gradients may be
 * NULL (zero). Materials are the first-layer view (M <= 16). */
typedef struct wc_adjoint {
    const float* grad_power_total;       /* [cells] */
    const float* grad_field_total;       /* [cells] c64 */
    const float* grad_component_power;   /* [4, cells] */
    const float* grad_component_field;   /* [4, cells] c64 */
    const float* grad_diffraction_power; /* [cells]: of sum_axis |diffraction_vector|^2 */
    const float* power_total;            /* forward outputs */
    const float* field_total;
    const float* component_field;
    const float* diffraction_vector;
    const int* topology;
    const int* path_count;
    int topology_capacity;
    float* partials;
} wc_adjoint;

/* Tangent of the radiomap along one direction of its parameters (forward
 * mode; reference ad_mode="jvp", the jvp companions of
 * witwin/channel/kernels/deterministic.py:156-183 and
 * native/channel/kernels/fields.cu:2183): d_eps_r / d_sigma_e [M] (first
 * layer of each material) and d_tx_pos [T, 3] may be NULL (zero). primal
 * carries the forward outputs and the stored topology (its grad_* and
 * partials fields are unused). Outputs: tangents of power_total [cells],
 * field_total [cells] c64, component_power [4, cells], component_field
 * [4, cells] c64 and (optional) the diffraction power sum_axis
 * |diffraction_vector|^2 [cells]. */
typedef struct wc_tangent_in {
    const float* d_eps_r;
    const float* d_sigma_e;
    float d_frequency;
    const float* d_tx_pos;
} wc_tangent_in;
typedef struct wc_tangent_out {
    float* d_power_total;
    float* d_field_total;
    float* d_component_power;
    float* d_component_field;
    float* d_diffraction_power;
} wc_tangent_out;
int wc_radiomap_jvp(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* primal,
                    const wc_tangent_in* tangent, const wc_tangent_out* out, void* stream);

int wc_vjp_partial_stride(int material_count);
int wc_vjp_partial_rows(int tile_count);
/* Deterministic reduction of the partial rows (float64 sums in a fixed
 * order: strided per-thread sums, then a fixed-shape tree; one workgroup per
 * output): grad_params [2M+1] = [dL/deps_r, dL/dsigma, dL/df],
 * grad_tx [tx_count, 3]. tiles are the (transmitter, first cell) pairs of the
 * launch. */
int wc_vjp_reduce(const float* partials, int tile_count, int stride, const int* tiles, int tx_count,
                  float* grad_params, float* grad_tx, void* stream);
int wc_radiomap_vjp(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint, const int* tiles,
                    int tile_count, void* stream);

/* One launch of wc_radiomap_vjp, for callers that time or schedule the
 * adjoint's launches apart: phase 1 (line of sight, reflection,
 * transmission paths) writes each tile's partial vector, phase 2
 * (diffraction paths, solves with diffraction only) adds to it, so phase 2
 * must follow phase 1 of the same adjoint. No reference counterpart. */
int wc_radiomap_vjp_phase(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint,
                          const int* tiles, int tile_count, int phase, void* stream);

/* Geometry part of the fixed-topology VJP (mesh vertices). Replaces the
 * reference's AD geometry re-solve, _reflection_geometry_ad (witwin/channel/
 * propagation/geometry.py:214-292) with RayD's EPC backward, and the
 * position/normal gradients of the field AD functions (witwin/channel/
 * kernels/fields.py:1283, 1560, 1827). Each stored path is recomputed with
 * its interaction geometry carrying tangents and the contracted derivatives
 * are ADDED (float64 atomics; the caller zeroes the buffers) to
 *   grad_group_plane [group_count, 4]: dL/d(n_x, n_y, n_z, c) of each group
 *     plane n . x = c (reflections and transmitted walls),
 *   grad_edge [edge_count, 8]: dL/d(origin xyz, frame rotation xyz, length,
 *     wedge index n) of each diffracting edge.
 * adjoint->partials is not used. The edge part is currently switched off:
 * solves with diffraction (and edges) return an error, grad_edge is not
 * written. */
int wc_radiomap_vjp_geometry(const wc_scene* scene, const wc_solve* solve, const wc_adjoint* adjoint,
                             const int* tiles, int tile_count, double* grad_group_plane, double* grad_edge,
                             void* stream);

/* Vertex gradient from the geometry accumulators: each group plane is the
 * plane of its representative face group_face[g] (faces [F, 3] vertex ids),
 * each edge a function of its vertices edge_vertices [E, 4] = (a, b, c0, c1;
 * c1 = -1 for a half-plane). contrib (workspace, (9 G + 12 E) floats)
 * receives the per-entity vertex contributions; grad_vertices [V, 3] their
 * sums per vertex over the CSR lists (csr_offset [V + 1], csr_slot: indices
 * of contribution triples, groups g*3 + k first, then 3 G + e*4 + k), in
 * list order (deterministic). No reference counterpart (the reference
 * scatters through RayD's backward). */
int wc_geometry_vertex_grad(const float* vertices, const int* faces, const int* group_face, int group_count,
                            const double* grad_group_plane, const int* edge_vertices, int edge_count,
                            const double* grad_edge, const int* csr_offset, const int* csr_slot, int vertex_count,
                            float* contrib, float* grad_vertices, void* stream);

/* ------------------------------------------------------------ path rows
 * Explicit path export (deterministic export_paths, path.solve,
 * consumer.evaluate). Pass 1 counts the valid paths of every cell; the
 * caller scans the counts in canonical (rx-major, tx-minor) pair order into
 * row_offset [cells]; pass 2 writes each cell's rows at row_offset[cell] in
 * discovery order, never more than the path_count [cells] of pass 1; wc_paths_canonical_order then sorts every pair segment by
 * (depth, component, primitive sequence, primitive, edge), stable, marks the
 * first strictly-shortest row of each identity group and returns the
 * permutation. Replaces the reference topology stages + canonical compact
 * (native/channel/kernels/evaluated_paths.cu:513-965, paths.cu:1808-2007). */
int wc_paths_count(const wc_scene* scene, const wc_solve* solve, int* path_count, void* stream);

typedef struct wc_path_rows {
    int width; /* sequence width (max depth, >= 1) */
    int* tx_id;
    int* rx_id;
    int* depth;
    int* component_id;
    int* primitive_id;
    int* edge_id;
    int* material_id;
    float* path_length_m;
    float* delay_s;
    float* path_gain;
    int* primitive_sequence;      /* [K, width] */
    int* material_sequence;       /* [K, width] */
    float* interaction_positions; /* [K, width, 3] */
    float* interaction_normals;   /* [K, width, 3] */
    float* path_field;            /* [K] c64 */
    float* coefficient;           /* [K] c64 */
    float* field_xyz;             /* [K, 3] c64 */
    float* arrival_direction;     /* [K, 3] */
    float* departure_direction;   /* [K, 3] */
} wc_path_rows;

int wc_paths_emit(const wc_scene* scene, const wc_solve* solve, const int* row_offset, const int* path_count,
                  const wc_path_rows* rows, void* stream);

int wc_paths_canonical_order(const wc_path_rows* rows, const int* pair_offset, int pair_count, int* permutation,
                             unsigned char* keep, void* stream);

/* Device-side canonical compaction (reference canonical compact,
 * native/channel/kernels/evaluated_paths.cu:355-384 (max_paths), 513-965;
 * pair offsets there come from the topology stages' scans, paths.cu:1808-2007).
 * Workspace of the three calls below for up to `cells` pairs. */
size_t wc_paths_compact_workspace_bytes(int cells);

/* Canonical pair order: pair_cell [cells] receives the cell (transmitter t's
 * receiver r, cells laid out by cell_offset [T + 1]) at each canonical position,
 * receiver-major then transmitter (a stable radix sort of r * T + t; max_rx
 * bounds r, and max_rx * T must fit 32 bits). Depends on the layout only. */
int wc_paths_pair_order(const int* cell_offset, int tx_count, int cells, int max_rx, int* pair_cell, void* workspace,
                        void* stream);

/* Row offsets of the export: pair_offset [cells + 1] = exclusive scan of
 * path_count [cells] taken in canonical order (path_count[pair_cell[i]]),
 * row_offset [cells] = the same offsets by cell (wc_paths_emit's input), and
 * *total (device int64) = the row count; the caller checks it fits int32. */
int wc_paths_row_offsets(const int* path_count, const int* pair_cell, int cells, int* row_offset, int* pair_offset,
                         long long* total, void* workspace, void* stream);

/* Canonical order, identity de-duplication and per-pair cap in one call:
 * permutation / keep as wc_paths_canonical_order; each pair then keeps its
 * identity winners in sorted order, at most max_paths_per_pair of them (< 0:
 * no cap); kept_offset [pair_count + 1] is the scan of the kept counts,
 * source_row [rows] lists the kept rows' indices in the emitted table in final
 * order (the caller gathers every field through it), *kept_total (device
 * int64) the kept row count and *dropped (device) the identity winners the cap
 * removed. */
int wc_paths_compact(const wc_path_rows* rows, const int* pair_offset, int pair_count, int max_paths_per_pair,
                     int* permutation, unsigned char* keep, int* kept_offset, int* source_row, long long* kept_total,
                     int* dropped, void* workspace, void* stream);

/* ------------------------------------------------------------ flat accumulation, reference arithmetic
 * deterministic_accumulate_flat / _fwd64 / _backward / _jvp of the reference
 * (witwin/channel/kernels/deterministic.py:23-390,
 * native/channel/kernels/deterministic.cu:829-1262): rows scatter into six
 * slots per (tx, rx) cell (component 0/1/2 -> 0/1/2, 5 -> 3, 6 -> 4
 * scattering, 3/4/7 -> 5 coupled); coherent cells square the summed field of
 * the field slots and add the scattering power, incoherent cells sum powers
 * with sqrt(P) as pseudo-field. Arithmetic in the row precision (dtype 0:
 * float32, 1: float64), per-slot sums in row order. Outputs: power_total
 * [T*R], field_total_real/imag [T*R], component_power and
 * component_field_real/imag [6, T*R].
 * wc_accumulate_rows_jvp: rows carry the tangents (any of the three may be
 * NULL = zero) and `forward` the primal outputs; writes the output tangents.
 * wc_accumulate_rows_backward: cotangents of the outputs (any NULL = zero)
 * -> per-row gradients of path_gain / field_real / field_imag. */
typedef struct wc_accum_rows {
    int count;
    int num_tx;
    int num_rx;
    int coherent;
    int scattering_coherent; /* scattering_combine_domain: 0 power, 1 coherent */
    int dtype;               /* 0 float32, 1 float64 */
    const unsigned char* valid;
    const int* tx_id;
    const int* rx_id;
    const int* component_id;
    const void* path_gain;
    const void* field_real;
    const void* field_imag;
} wc_accum_rows;
typedef struct wc_accum_out {
    void* power_total;
    void* field_total_real;
    void* field_total_imag;
    void* component_power;
    void* component_field_real;
    void* component_field_imag;
} wc_accum_out;
size_t wc_accumulate_rows_workspace_bytes(int count, int num_tx, int num_rx);
int wc_accumulate_rows(const wc_accum_rows* rows, const wc_accum_out* out, void* workspace, void* stream);
int wc_accumulate_rows_jvp(const wc_accum_rows* tangents, const wc_accum_out* forward, const wc_accum_out* out,
                           void* workspace, void* stream);
int wc_accumulate_rows_backward(const wc_accum_rows* rows, const wc_accum_out* forward, const wc_accum_out* grad_out,
                                void* grad_path_gain, void* grad_field_real, void* grad_field_imag, void* stream);

/* ------------------------------------------------------------ chain cones
 * Receiver cone of every chain: a conservative bound, as planes, of the
 * receivers whose back-projection can land inside the convex hull of every
 * group of the chain (the hull of group g: hull_points [H, 3] float64 rows
 * hull_offset[g] .. hull_offset[g + 1], in the group's plane). groups [K, 5],
 * images [K, 5, 3] and depth [K] as in the chain plan; cones [K,
 * WC_CONE_FLOATS]. Lets the chain culls drop (tile, chain) pairs that no
 * receiver of the tile can validate. No reference counterpart (the reference
 * tests every candidate for every receiver). */
#define WC_CONE_FLOATS 48
int wc_chain_cones(const wc_scene* scene, const int* hull_offset, const double* hull_points, const int* groups,
                   const double* images, const int* depth, int count, float* cones, void* stream);

/* ------------------------------------------------------------ chain planning
 * Candidate reflection chains from discovery rays, on the device (reference
 * witwin/channel/propagation/reflection.py:186-205, 541-569: the distinct
 * group-sequence prefixes of each transmitter's rays, per depth, sorted).
 * rays [T, R, max_depth] are wc_discover_chains' output (-1 = escaped).
 * Per depth d = 1, 2, ...: wc_plan_level sorts the rays' d-prefixes and
 * counts the distinct ones into *prefix_count (device int64); the caller
 * allocates that many entries per array and wc_plan_assign numbers them in
 * (transmitter, lexicographic) order: prefix_parent (id at depth d - 1, or
 * the transmitter for d = 1), prefix_group (d-th group), prefix_tx, and adds
 * them to counts [T, WC_PLAN_DEPTHS] (zeroed by the d = 1 level call).
 * parent_tx is depth d - 1's prefix_tx (unused for d = 1). The workspace
 * (wc_plan_workspace_bytes(T, R)) carries the rays' prefix ids from one
 * depth to the next, so the depths run in order on one workspace.
 * wc_plan_table writes depth d's prefixes as chain rows (groups [K, 5],
 * padded with -1, their transmitter and depth) at row row_start[tx] + id;
 * wc_plan_exhaustive writes one depth's exhaustive sequence list
 * [count, d] for every transmitter at rows row_start[tx] + j. */
#define WC_PLAN_DEPTHS 5
typedef struct wc_plan_levels {
    const int* prefix_parent[WC_PLAN_DEPTHS];
    const int* prefix_group[WC_PLAN_DEPTHS];
    const int* prefix_tx[WC_PLAN_DEPTHS];
} wc_plan_levels;

size_t wc_plan_workspace_bytes(int tx_count, int ray_count);
int wc_plan_level(const int* rays, int tx_count, int ray_count, int max_depth, int d, int group_count, int* counts,
                  long long* prefix_count, void* workspace, void* stream);
int wc_plan_assign(const int* rays, int tx_count, int ray_count, int max_depth, int d, int group_count,
                   const int* parent_tx, int* prefix_parent, int* prefix_group, int* prefix_tx, int* counts,
                   void* workspace, void* stream);
int wc_plan_table(const wc_plan_levels* levels, int d, int prefix_count, const int* row_start, int* groups,
                  int* chain_tx, int* chain_depth, void* stream);
int wc_plan_exhaustive(const int* sequences, int count, int d, int tx_count, const int* row_start, int* groups,
                       int* chain_tx, int* chain_depth, void* stream);

/* ------------------------------------------------------------ channel frequency response
 * CFR of padded per-pair path sets, replacing PathResult.cfr
 * (witwin/channel/path.py:646-656) and BeamformedPathResult.cfr (:824-832):
 *   out[b, s, k] = sum_p valid[b,p] a[b,p,s] exp(-j 2 pi f_k (tau[b,p] - tau0[b]))
 * a [pairs, max_paths, time_steps] c64, tau [pairs, max_paths] f32 (s),
 * valid [pairs, max_paths] u8, out [pairs, time_steps, freq_count] c64.
 * tau0[b] = min over valid paths of pair b when normalize_delays != 0, else 0
 * (tau - tau0 rounded to float32 as the reference does). Frequencies are the
 * list `frequencies` [freq_count] f64 (Hz) or, when it is NULL, the uniform
 * grid f0 + k df. The phase argument f (tau - tau0) is reduced modulo one
 * cycle in float64; path sums run in float32 in path order. */
int wc_paths_cfr(const float* a, const float* tau, const unsigned char* valid, int pairs, int max_paths,
                 int time_steps, const double* frequencies, double f0, double df, int freq_count,
                 int normalize_delays, float* out, void* stream);

/* ------------------------------------------------------------ row field operators
 * Per-row field kernels with the reference facade signatures
 * (witwin/channel/kernels/fields.py field_free_space :110,
 *  field_project_complex3 :149, field_reflection_sequence :180;
 *  native/channel/kernels/fields.cu:36-204). Outputs per row: field_vector
 * [N,3] c64, coefficient [N] c64, path_field [N] c64, path_gain [N],
 * path_length [N], delay [N], direction [N,3]. The frequency is float64: the
 * phase -k L is formed in float64 from it, the Fresnel terms in float32. */
int wc_field_free_space(int count, const float* source, const float* target, const float* tx_power,
                        const float* tx_polarization, const float* rx_polarization, double frequency_hz,
                        float* field_vector, float* coefficient, float* path_field, float* path_gain,
                        float* path_length, float* delay, float* direction, void* stream);

/* Line-of-sight export of every (transmitter, receiver) pair, replacing
 * path_los_export (reference kernels/topology.py:84, los.cu:74-125): pair
 * index rx * T + tx; path_length = max(|tx - rx|, 1e-6), delay = length / c,
 * gain = P (lambda / (4 pi d))^2 (|p|^2 - (p . k)^2); path_gain_matrix [T, R].
 * No visibility test (the caller gates the pairs). */
int wc_path_los_export(int tx_count, int rx_count, const float* tx_positions, const float* tx_power,
                       const float* rx_positions, const float* tx_polarizations, double frequency_hz, int* tx_id,
                       int* rx_id, float* path_length, float* delay, float* path_gain, float* path_gain_matrix,
                       void* stream);

int wc_field_project_complex3(int count, const float* field_vector, const float* direction,
                              const float* rx_polarization, float* coefficient, float* path_gain, void* stream);

int wc_field_reflection_sequence(int count, int depth, const float* source, const float* target,
                                 const float* interaction_positions, const float* interaction_normals,
                                 const float* tx_power, const float* tx_polarization, const float* rx_polarization,
                                 const float* eps_r, const float* sigma_e, const float* mu_r, const float* gain,
                                 const float* thickness, double frequency_hz, float* field_vector, float* coefficient,
                                 float* path_field, float* path_gain, float* path_length, float* delay,
                                 float* direction, void* stream);

/* ------------------------------------------------------------ coupled reflection-diffraction paths
 * The radiomap's coupled slot (reference interactions/coupled.py:93-230,
 * native/channel/kernels/coupled.cu:2279-2710; deterministic.py:120-127
 * coupled_paths): per (transmitter, receiver) cell every (coplanar group,
 * edge) candidate as TX -> reflect -> diffract -> RX (component 3) and as
 * TX -> diffract -> reflect -> RX (component 4), then every ordered edge pair
 * e1 != e2 as TX -> diffract -> diffract -> RX (component 7). Stationary points
 * in float64 (image method, Keller points, a bisection along edge 2 for
 * double diffraction), strictly inside the edges and the reflecting face,
 * every leg unoccluded; fields composed from the thin-slab Jones reflection and
 * the first-order UTD wedge coefficient. kinds: 1 R->D, 2 D->R, 4 D->D.
 * Cells [cell_begin, cell_end) are evaluated; the outputs are indexed by cell:
 * field [cells] c64 = sum of sqrt(P) x receiver-projected path fields,
 * power [cells] = sum of path gains, count [cells] = paths found (sums in
 * float64 in candidate order). The workspace holds 16 B per (cell, candidate)
 * of the range: wc_coupled_workspace_bytes(cell_end - cell_begin, G, E, kinds). */
typedef struct wc_coupled_out {
    float* field;
    float* power;
    int* count;
    int* count_dd; /* [cells] double-diffraction paths among count, or NULL */
} wc_coupled_out;

long long wc_coupled_candidates_per_cell(int group_count, int edge_count, int kinds);
size_t wc_coupled_workspace_bytes(int cells, int group_count, int edge_count, int kinds);
int wc_coupled_paths(const wc_scene* scene, const wc_solve* solve, int cell_begin, int cell_end, int kinds,
                     const wc_coupled_out* out, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------ row operators: transmission, layer stacks, AD
 * Layered-material tables in CSR form (reference kernels/materials.py:42-70
 * _validate_layer_csr): material m owns layers layer_offset[m] ..
 * layer_offset[m] + layer_count[m] - 1, ordered entry -> exit, at most 8
 * layers. A row that names a material outside the tables, or one whose span
 * is negative, longer than 8 or past layer_total, gives zeros. */
typedef struct wc_layer_tables {
    const int* layer_offset;      /* [material_count] */
    const int* layer_count;       /* [material_count] */
    const float* layer_thickness; /* [layer_total] metres */
    const float* layer_eps;       /* [layer_total] relative permittivity */
    const float* layer_sigma;     /* [layer_total] conductivity S/m */
    const float* layer_mu;        /* [layer_total] relative permeability */
    int material_count;
    int layer_total;
} wc_layer_tables;

/* The enumerated-field outputs of one row set (the reference's shared
 * result schema, kernels/fields.py:85-107): field_vector [N,3] c64,
 * coefficient / path_field [N] c64, path_gain, path_length_m, delay_s [N],
 * direction [N,3]. As tangents (the _jvp entry points) they receive
 * d(output); as cotangents (wc_field_rows_grad) NULL means zero. */
typedef struct wc_field_rows_out {
    float* field_vector;
    float* coefficient;
    float* path_field;
    float* path_gain;
    float* path_length_m;
    float* delay_s;
    float* direction;
} wc_field_rows_out;

typedef struct wc_field_rows_grad {
    const float* field_vector;
    const float* coefficient;
    const float* path_field;
    const float* path_gain;
    const float* path_length_m;
    const float* delay_s;
    const float* direction;
} wc_field_rows_grad;

/* Bytes of the workspace the _vjp entry points take: per-layer float64 sums
 * and per-row frequency partials, reduced in a fixed order. */
size_t wc_field_rows_workspace_bytes(int rows, int layers);

/* Free-space rows, forward and reverse mode (reference kernels/fields.py:551
 * field_free_space_backward, :583 field_free_space_jvp). Tangent pointers
 * may be NULL (zero). grad_source / grad_target may be NULL (not needed);
 * grad_frequency is one float. */
int wc_field_free_space_jvp(int count, const float* source, const float* target, const float* tx_power,
                            const float* tx_polarization, const float* rx_polarization, double frequency_hz,
                            float tangent_frequency, const float* tangent_source, const float* tangent_target,
                            const wc_field_rows_out* tangents, void* stream);
int wc_field_free_space_vjp(int count, const float* source, const float* target, const float* tx_power,
                            const float* tx_polarization, const float* rx_polarization, double frequency_hz,
                            const wc_field_rows_grad* grads, float* grad_source, float* grad_target,
                            float* grad_frequency, void* workspace, void* stream);

/* Reflection-sequence rows (reference kernels/fields.py:180, :607 backward,
 * :665 jvp): depth D bounces, per-interaction arrays [N, D] (positions and
 * normals [N, D, 3]). */
typedef struct wc_reflection_rows {
    int count;
    int depth;
    double frequency_hz;
    const float* source;
    const float* target;
    const float* interaction_positions;
    const float* interaction_normals;
    const float* tx_power;
    const float* tx_polarization;
    const float* rx_polarization;
    const float* eps_r;
    const float* sigma_e;
    const float* mu_r;
    const float* gain;
    const float* thickness;
} wc_reflection_rows;

typedef struct wc_reflection_rows_gradout {
    float* grad_source;                /* each may be NULL (not needed) */
    float* grad_target;
    float* grad_interaction_positions;
    float* grad_interaction_normals;
    float* grad_eps_r;
    float* grad_sigma_e;
    float* grad_gain;
    float* grad_thickness;
    float* grad_frequency;             /* one float, required */
} wc_reflection_rows_gradout;

/* tangents: the same layout as rows; each tangent pointer may be NULL (zero);
 * its count, depth, frequency, tx_power, polarizations and mu_r are unused. */
int wc_field_reflection_sequence_jvp(const wc_reflection_rows* rows, const wc_reflection_rows* tangents,
                                     float tangent_frequency, const wc_field_rows_out* out, void* stream);
int wc_field_reflection_sequence_vjp(const wc_reflection_rows* rows, const wc_field_rows_grad* grads,
                                     const wc_reflection_rows_gradout* out, void* workspace, void* stream);

/* Straight transmission through thin layered walls (reference
 * kernels/fields.py:245 field_transmission_sequence, :710 backward, :771
 * jvp): per valid interaction the layered t_TE / t_TM on the wall's s/p
 * basis, carrier phase over L - sum(thickness cos theta), amplitude
 * 1 / (2 k L). Rows with path_valid == 0, or with a valid interaction whose
 * material is outside the tables, give zeros. interaction_positions of the
 * reference signature do not enter the field and are not passed. */
typedef struct wc_transmission_rows {
    int count;
    int depth; /* 1..8 */
    double frequency_hz;
    const unsigned char* path_valid;        /* [N] */
    const unsigned char* interaction_valid; /* [N, D] */
    const int* interaction_material_id;     /* [N, D] */
    const float* source;
    const float* target;
    const float* interaction_normals; /* [N, D, 3] */
    const float* tx_power;
    const float* tx_polarization;
    const float* rx_polarization;
    wc_layer_tables layers;
} wc_transmission_rows;

int wc_field_transmission_sequence(const wc_transmission_rows* rows, const wc_field_rows_out* out, void* stream);
int wc_field_transmission_sequence_jvp(const wc_transmission_rows* rows, const float* tangent_source,
                                       const float* tangent_target, const float* tangent_normals,
                                       const float* tangent_layer_thickness, const float* tangent_layer_eps_r,
                                       const float* tangent_layer_sigma_e, float tangent_frequency,
                                       const wc_field_rows_out* out, void* stream);
int wc_field_transmission_sequence_vjp(const wc_transmission_rows* rows, const wc_field_rows_grad* grads,
                                       float* grad_source, float* grad_target, float* grad_normals,
                                       float* grad_layer_thickness, float* grad_layer_eps_r,
                                       float* grad_layer_sigma_e, float* grad_frequency, void* workspace,
                                       void* stream);

/* Layer stack between vacuum half-spaces (reference kernels/materials.py:151
 * em_layer_stack_eval, :199 em_layer_stack_backward, :235
 * em_layer_stack_jvp): amplitude r, t per polarization (backward Rouard
 * recursion) and R = |r|^2, T = |t|^2. A zero-layer material gives r = 0,
 * t = 1; a material id outside the tables gives zeros. As cotangents of the
 * _vjp entry point the twelve pointers are read only, NULL meaning zero. */
typedef struct wc_layer_stack_out {
    float* r_te_real;
    float* r_te_imag;
    float* r_tm_real;
    float* r_tm_imag;
    float* t_te_real;
    float* t_te_imag;
    float* t_tm_real;
    float* t_tm_imag;
    float* cap_R_te;
    float* cap_R_tm;
    float* cap_T_te;
    float* cap_T_tm;
} wc_layer_stack_out;

int wc_em_layer_stack_eval(int count, const float* cos_theta, const int* material_id, const wc_layer_tables* layers,
                           double frequency_hz, const wc_layer_stack_out* out, void* stream);
int wc_em_layer_stack_jvp(int count, const float* cos_theta, const int* material_id, const wc_layer_tables* layers,
                          double frequency_hz, const float* tangent_cos_theta, const float* tangent_layer_thickness,
                          const float* tangent_layer_eps_r, const float* tangent_layer_sigma_e,
                          float tangent_frequency, const wc_layer_stack_out* out, void* stream);
int wc_em_layer_stack_vjp(int count, const float* cos_theta, const int* material_id, const wc_layer_tables* layers,
                          double frequency_hz, const wc_layer_stack_out* grads, float* grad_cos_theta,
                          float* grad_layer_thickness, float* grad_layer_eps_r, float* grad_layer_sigma_e,
                          float* grad_frequency, void* workspace, void* stream);

/* ------------------------------------------------------------ accumulation
 * Deterministic flat accumulation (reference deterministic.cu:829-970,
 * witwin/channel/kernels/deterministic.py:23-102): rows are ordered by cell
 * with a stable radix sort and each cell is reduced in row order in float64,
 * so results are bit-reproducible. Slot map: component 0/1/2 -> 0/1/2,
 * 5 -> 3, 6 -> 4, 3/4/7 -> 5. Outputs: component_power [6, T*R],
 * component_field [6, T*R] c64, power_total [T*R], field_total [T*R] c64.
 * workspace: wc_accumulate_workspace_bytes(count, num_tx, num_rx) bytes. */
size_t wc_accumulate_workspace_bytes(int count, int num_tx, int num_rx);
int wc_accumulate_flat(int count, const unsigned char* valid, const int* tx_id, const int* rx_id,
                       const int* component_id, const float* path_gain, const float* field_real,
                       const float* field_imag, int num_tx, int num_rx, int coherent, float* component_power,
                       float* component_field, float* power_total, float* field_total, void* workspace,
                       size_t workspace_bytes, void* stream);

/* Allocation-free variant of wc_accumulate_flat. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as wc_accumulate_flat without the workspace: the
 * sort keys, row ids, histogram and segment starts live in a fixed
 * device-resident scratch array of wc_accumulate_noalloc_max_bytes() bytes
 * (64 MiB, one per GPU, held until the process ends). Calls whose
 * wc_accumulate_workspace_bytes(count, num_tx, num_rx) exceeds that fail with
 * a message and launch nothing. It calls no allocator. Every call on one GPU
 * shares that GPU's scratch: calls that overlap in time on one GPU (two
 * streams, or two host threads) race and must be serialized by the caller. */
size_t wc_accumulate_noalloc_max_bytes(void);
int wc_accumulate_flat_noalloc(int count, const unsigned char* valid, const int* tx_id, const int* rx_id,
                               const int* component_id, const float* path_gain, const float* field_real,
                               const float* field_imag, int num_tx, int num_rx, int coherent,
                               float* component_power, float* component_field, float* power_total,
                               float* field_total, void* stream);

/* ------------------------------------------------------------ visibility
 * Segment visibility (reference kernels/geometry.py rayd_visibility_forward
 * :380 -> RayD). visible[i] = 1 when segment source[i] -> target[i] crosses
 * no triangle (end points excluded by a 1e-4 m clearance). */
int wc_visibility(const wc_scene* scene, int count, const float* source, const float* target,
                  unsigned char* visible, void* stream);

/* ------------------------------------------------------------ fixed-topology replay
 * Replays frozen path rows at new endpoints (consumer reevaluate). Replaces
 * the per-bucket replay of witwin/channel/propagation/consumer.py:3029-3096
 * (_los_inputs :2770 with the RayD visibility gate, _reflection_inputs :2734
 * -> reflection_epc_paths, propagation/geometry.py:214, then
 * field_free_space / field_reflection_sequence). component 0 (line of sight,
 * depth 0) re-tests its segment; component 1 (reflection, depth 1..width)
 * re-solves the image method over the planes of its frozen faces' groups and
 * re-checks faces and legs. Rows that no longer exist get valid = 0 and zero
 * fields; other components are invalid. One launch covers every depth. */
typedef struct wc_fixed_rows {
    int count;
    int width;                    /* interaction slots per row */
    const int* component;         /* [N] */
    const int* depth;             /* [N] */
    const int* faces;             /* [N, width] frozen face ids */
    const float* source;          /* [N, 3] */
    const float* target;          /* [N, 3] */
    const float* tx_power;        /* [N] W */
    const float* tx_polarization; /* [N, 3] */
    const float* rx_polarization; /* [N, 3] */
    unsigned char* valid;         /* [N] out */
    float* path_field;            /* [N] c64 out: coefficient * sqrt(power) */
    float* field_xyz;             /* [N, 3] c64 out: unit-excitation field at the target */
    float* path_length;           /* [N] out, m */
    float* delay;                 /* [N] out, s */
    float* direction;             /* [N, 3] out: arrival direction */
    float* positions;             /* [N, width, 3] out (zeros past the depth or when invalid) */
    float* normals;               /* [N, width, 3] out */
} wc_fixed_rows;

int wc_fixed_rows_evaluate(const wc_scene* scene, const wc_fixed_rows* rows, double frequency_hz, void* stream);

/* ------------------------------------------------------------ discovery
 * Transmitter-launched specular discovery rays (reference
 * native/channel/rayd/reflection.cpp:9-21, witwin/channel/interactions/
 * reflection.py:541-569): ray_count directions on a Fibonacci sphere from
 * each transmitter, traced with closest hits and specular continuation up to
 * max_depth bounces. chains [tx_count, ray_count, max_depth] i32 receives the
 * coplanar group hit at each bounce (-1 after the ray escapes). */
int wc_discover_chains(const wc_scene* scene, const float* tx_pos, int tx_count, int ray_count, int max_depth,
                       int* chains, void* stream);

/* Image points of candidate chains: images[k, j] = source of the chain's
 * transmitter mirrored successively across groups 0..j (float64). */
int wc_chain_images(const wc_scene* scene, const float* tx_pos, const int* chain_tx, const int* chain_groups,
                    const int* chain_depth, int chain_count, double* images, void* stream);

/* ------------------------------------------------------------ diagnostics
 * Element-wise evaluation of the device special functions the UTD
 * coefficient uses: kind 0 = Kouyoumjian-Pathak transition function F(x),
 * kind 1 = complex Fresnel integral C(x) + j S(x); x [n] f64, out [n] c128
 * (interleaved re, im). No reference counterpart: it exposes the kernels'
 * own evaluation (float64, Chebyshev/asymptotic Fresnel auxiliary functions)
 * to accuracy tests. */
int wc_utd_special_eval(int kind, const double* x, int n, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WCMI_H */
