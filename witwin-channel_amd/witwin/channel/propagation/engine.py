"""Host orchestration of the MI355X propagation engine.

This is the counterpart of the reference stage driver
``propagation.enumerated.evaluate_enumerated_paths``
(witwin/channel/propagation/enumerated.py:549-837). The reference runs one
Python loop per transmitter x receiver chunk x depth, materializes every
candidate row, compacts, sorts and only then evaluates fields. Here the
whole (transmitter, cell) product is one fused kernel launch:

1. candidate chains per transmitter (exhaustive face-group sequences when
   their count is small, transmitter-launched discovery rays otherwise; the
   thresholds are the reference's: reflection.py:30-34, 120, 186-205),
2. double-precision image points of every chain (one small kernel),
3. ``wc_radiomap_forward`` (fused discovery + field + float64 coherent sum per
   cell), or ``wc_paths_count`` / ``wc_paths_emit`` /
   ``wc_paths_canonical_order`` when explicit path rows are requested.

Host<->device synchronization happens once per row export (the row count),
like the reference's single count read (docs/dev/deterministic/README.md:30-36).
"""

from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from witwin.channel import _native
from witwin.channel.propagation import planning
from witwin.channel.scene.compiler import CompiledScene, DeviceScene

MAX_DEPTH = 5
EXHAUSTIVE_GROUP_LIMIT = 4096        # depth-1 exhaustive when groups <= this
EXHAUSTIVE_SEQUENCE_LIMIT = 100_000  # depth >= 2 exhaustive when G (G-1)^(d-1) <= this
DISCOVERY_RAY_COUNT = 262_144
TILE = 64
# lanes sharing one cell's candidates in the radiomap and adjoint kernels.
# More lanes give more wavefronts per map, which pays only once the kernels'
# register footprint lets several wavefronts share a SIMD; measured on the
# 256 x 256 urban canyon: L=1 17.4 ms, L=2 17.9, L=4 18.9, L=8 19.7, L=16 22.1.
LANES_PER_CELL = 1


@dataclass
class ChainPlan:
    offset: torch.Tensor        # [T*5+1] int32
    groups: torch.Tensor        # [K,5] int32
    images: torch.Tensor        # [K,5,3] float64
    depth: torch.Tensor         # [K] int32
    tx: torch.Tensor            # [K] int32
    offset_host: np.ndarray
    discovered_depths: tuple[int, ...]
    _groups_host: np.ndarray | None = None
    cones: torch.Tensor | None = None   # [K, CONE_FLOATS] receiver cones (chain_cones)

    @property
    def count(self) -> int:
        return int(self.offset_host[-1])

    @property
    def groups_host(self) -> np.ndarray:
        """[K, 5] chain groups on the host (copied on first use)."""
        if self._groups_host is None:
            k = self.count
            self._groups_host = self.groups[:k].cpu().numpy() if k else np.zeros((0, MAX_DEPTH), np.int32)
        return self._groups_host


exhaustive_sequences = planning.exhaustive_sequences


def is_exhaustive(group_count: int, depth: int) -> bool:
    if depth == 1:
        return group_count <= EXHAUSTIVE_GROUP_LIMIT
    return group_count * (group_count - 1) ** (depth - 1) <= EXHAUSTIVE_SEQUENCE_LIMIT


def discover(dscene: DeviceScene, tx_pos: torch.Tensor, max_depth: int, ray_count: int) -> np.ndarray:
    """[T, ray_count, max_depth] int32 group hit per bounce (-1 = escaped)."""
    t = int(tx_pos.shape[0])
    chains = torch.empty((t, ray_count, max_depth), dtype=torch.int32, device=tx_pos.device)
    _native.call("wc_discover_chains", dscene.ref(), _native.ptr(tx_pos), t, ray_count, max_depth,
                 _native.ptr(chains), _native.stream_handle(tx_pos.device))
    return chains.cpu().numpy()


def unique_prefixes(rows: np.ndarray, depth: int) -> np.ndarray:
    pref = rows[:, :depth]
    pref = pref[(pref >= 0).all(axis=1)]
    if pref.shape[0] == 0:
        return np.zeros((0, depth), np.int32)
    return np.unique(pref, axis=0).astype(np.int32)


def chain_sequences(compiled: CompiledScene, dscene: DeviceScene, tx_pos: torch.Tensor, max_depth: int, *,
                    reflection: bool, ray_count: int = DISCOVERY_RAY_COUNT,
                    discovered: np.ndarray | None = None) -> list[list[np.ndarray]]:
    """Candidate group sequences of every transmitter: ``[T][d - 1]`` int32
    arrays [K_d, d] for depths 1..5 (empty above ``max_depth``), each sorted
    lexicographically — exhaustive when the count is small, otherwise the
    distinct prefixes of the transmitter's discovery rays (reference
    reflection.py:186-205, 541-569)."""
    if discovered is None:
        seqs = chain_sequences_device(compiled, dscene, tx_pos, max_depth, reflection=reflection, ray_count=ray_count)
        return [seqs.host(t) for t in range(len(seqs))]
    t = int(tx_pos.shape[0])
    g = compiled.group_count
    depths = range(1, max_depth + 1) if reflection and g > 0 else range(0)
    exhaustive_cache: dict[int, np.ndarray] = {}
    out = []
    for ti in range(t):
        per = []
        for d in range(1, MAX_DEPTH + 1):
            seqs = np.zeros((0, d), np.int32)
            if d in depths:
                if not is_exhaustive(g, d):
                    seqs = unique_prefixes(discovered[ti], d)
                else:
                    if d not in exhaustive_cache:
                        exhaustive_cache[d] = exhaustive_sequences(g, d)
                    seqs = exhaustive_cache[d]
            per.append(seqs)
        out.append(per)
    return out


def plan_from_sequences(dscene: DeviceScene, tx_pos: torch.Tensor, sequences: list[list[np.ndarray]],
                        discovered_depths: tuple[int, ...] = ()) -> ChainPlan:
    """ChainPlan of the transmitters ``tx_pos`` [T, 3] from their candidate
    sequences (``chain_sequences``): chains CSR by (transmitter, depth) with
    their double-precision image points."""
    device = tx_pos.device
    t = int(tx_pos.shape[0])
    blocks = []
    offsets = np.zeros(t * MAX_DEPTH + 1, dtype=np.int32)
    cursor = 0
    for ti in range(t):
        for d in range(1, MAX_DEPTH + 1):
            seqs = sequences[ti][d - 1]
            padded = np.full((seqs.shape[0], MAX_DEPTH), -1, dtype=np.int32)
            padded[:, :d] = seqs
            blocks.append((padded, ti, d))
            cursor += seqs.shape[0]
            offsets[ti * MAX_DEPTH + d] = cursor
    groups = np.concatenate([b[0] for b in blocks], axis=0) if blocks else np.zeros((0, MAX_DEPTH), np.int32)
    chain_tx = np.concatenate([np.full(b[0].shape[0], b[1], np.int32) for b in blocks]) if blocks else \
        np.zeros(0, np.int32)
    chain_depth = np.concatenate([np.full(b[0].shape[0], b[2], np.int32) for b in blocks]) if blocks else \
        np.zeros(0, np.int32)
    k = groups.shape[0]
    groups_t = torch.as_tensor(groups).to(device)
    tx_t = torch.as_tensor(chain_tx).to(device)
    depth_t = torch.as_tensor(chain_depth).to(device)
    images = torch.zeros((max(k, 1), MAX_DEPTH, 3), dtype=torch.float64, device=device)
    if k:
        _native.call("wc_chain_images", dscene.ref(), _native.ptr(tx_pos), _native.ptr(tx_t), _native.ptr(groups_t),
                     _native.ptr(depth_t), k, _native.ptr(images), _native.stream_handle(device))
    return ChainPlan(
        offset=torch.as_tensor(offsets).to(device), groups=groups_t if k else torch.full((1, MAX_DEPTH), -1,
                                                                                         dtype=torch.int32,
                                                                                         device=device),
        images=images, depth=depth_t, tx=tx_t, offset_host=offsets, discovered_depths=tuple(discovered_depths),
        _groups_host=groups,
    )


CONE_FLOATS = 48  # WC_CONE_FLOATS


def chain_cones(dscene: DeviceScene, chains: ChainPlan) -> torch.Tensor | None:
    """Receiver cones of the plan's chains (wc_chain_cones over the scene's
    group hulls), computed on first use; None without chains."""
    k = chains.count
    if k == 0:
        return None
    if chains.cones is None:
        hull_offset, hull_points = dscene.hulls()
        cones = torch.empty((k, CONE_FLOATS), dtype=torch.float32, device=dscene.device)
        _native.call("wc_chain_cones", dscene.ref(), _native.ptr(hull_offset), _native.ptr(hull_points),
                     _native.ptr(chains.groups), _native.ptr(chains.images), _native.ptr(chains.depth), k,
                     _native.ptr(cones), _native.stream_handle(dscene.device))
        chains.cones = cones
    return chains.cones


def refresh_chain_geometry(dscene: DeviceScene, inputs: "SolveInputs") -> None:
    """Recompute the chains' image points (transmitters or vertices moved) and,
    when the solve uses them, their receiver cones."""
    chains = inputs.chains
    if not chains.count:
        return
    _native.call("wc_chain_images", dscene.ref(), _native.ptr(inputs.tx_pos), _native.ptr(chains.tx),
                 _native.ptr(chains.groups), _native.ptr(chains.depth), chains.count, _native.ptr(chains.images),
                 _native.stream_handle(dscene.device))
    if inputs.view.chain_cones:
        chains.cones = None
        inputs.view.chain_cones = _native.ptr(chain_cones(dscene, chains))


def chain_sequences_device(compiled: CompiledScene, dscene: DeviceScene, tx_pos: torch.Tensor, max_depth: int, *,
                           reflection: bool, ray_count: int = DISCOVERY_RAY_COUNT) -> "planning.DeviceSequences":
    """Candidate group sequences of every transmitter, planned on the device
    (discovery, per-depth distinct prefixes, chain table; see
    propagation/planning.py): the same sets and order as ``chain_sequences``."""
    g = compiled.group_count
    return planning.device_sequences(g, dscene, tx_pos, max_depth, reflection=reflection,
                                     exhaustive=lambda d: is_exhaustive(g, d), ray_count=ray_count)


def plan_from_device(dscene: DeviceScene, tx_pos: torch.Tensor, sequences: "planning.DeviceSequences") -> ChainPlan:
    """ChainPlan of the transmitters ``tx_pos`` [T, 3] from their device
    sequences (``chain_sequences_device`` or its ``subset``): the table is used
    in place and only the image points are computed."""
    device = tx_pos.device
    k = sequences.count
    images = torch.zeros((max(k, 1), MAX_DEPTH, 3), dtype=torch.float64, device=device)
    if k:
        _native.call("wc_chain_images", dscene.ref(), _native.ptr(tx_pos), _native.ptr(sequences.chain_tx),
                     _native.ptr(sequences.groups), _native.ptr(sequences.chain_depth), k, _native.ptr(images),
                     _native.stream_handle(device))
    offsets = sequences.offset_host.astype(np.int32)
    return ChainPlan(offset=torch.as_tensor(offsets).to(device), groups=sequences.groups, images=images,
                     depth=sequences.chain_depth, tx=sequences.chain_tx, offset_host=offsets,
                     discovered_depths=tuple(sequences.discovered_depths))


def plan_chains(compiled: CompiledScene, dscene: DeviceScene, tx_pos: torch.Tensor, max_depth: int, *,
                reflection: bool, ray_count: int = DISCOVERY_RAY_COUNT,
                discovered: np.ndarray | None = None) -> ChainPlan:
    """Candidate reflection chains of every transmitter (sorted by depth, then
    lexicographically by group sequence)."""
    if discovered is None:
        return plan_from_device(dscene, tx_pos, chain_sequences_device(compiled, dscene, tx_pos, max_depth,
                                                                       reflection=reflection, ray_count=ray_count))
    g = compiled.group_count
    depths = range(1, max_depth + 1) if reflection and g > 0 else range(0)
    needs_rays = tuple(d for d in depths if not is_exhaustive(g, d))
    seqs = chain_sequences(compiled, dscene, tx_pos, max_depth, reflection=reflection, ray_count=ray_count,
                           discovered=discovered)
    return plan_from_sequences(dscene, tx_pos, seqs, needs_rays)


@dataclass
class SolveInputs:
    """Device-resident endpoint tables of one solve and the wc_solve view."""

    tx_pos: torch.Tensor
    tx_pol: torch.Tensor
    tx_power: torch.Tensor
    rx_pos: torch.Tensor
    rx_pol: torch.Tensor
    cell_offset: torch.Tensor
    rx_shared: bool
    chains: ChainPlan
    view: _native.WcSolve
    tiles: torch.Tensor
    tile_count: int
    rx_count: np.ndarray   # receivers per transmitter (host)
    cell_order: torch.Tensor | None = None  # evaluation position -> cell
    pair_cell: torch.Tensor | None = None   # canonical pair position -> cell (built on first export)

    @property
    def total_cells(self) -> int:
        return int(self.view.total_cells)


def make_tiles(cell_offset: np.ndarray, tile_cells: int = TILE) -> np.ndarray:
    """Wavefront work tiles (tx, first cell) of ``tile_cells`` cells that never
    straddle two transmitters."""
    bounds = np.asarray(cell_offset, dtype=np.int64)
    parts = [np.stack([np.full(len(r), t, np.int64), r], axis=1)
             for t in range(bounds.shape[0] - 1)
             for r in [np.arange(bounds[t], bounds[t + 1], tile_cells, dtype=np.int64)] if len(r)]
    if not parts:
        return np.zeros((0, 2), dtype=np.int32)
    return np.concatenate(parts, axis=0).astype(np.int32)


SUPER_TILES = 16  # tiles per supertile of the chain-validity cull


def super_tables(tiles: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Supertiles of a tile list (wc_solve.super_first_tile / tile_super):
    runs of at most ``SUPER_TILES`` consecutive tiles of one transmitter.
    Returns ([S + 1] first tile of each supertile, [tiles] supertile of each
    tile)."""
    tiles = np.asarray(tiles).reshape(-1, 2)
    n = tiles.shape[0]
    if n == 0:
        return np.zeros(1, np.int32), np.zeros(0, np.int32)
    tx = tiles[:, 0].astype(np.int64)
    start = np.ones(n, dtype=bool)
    start[1:] = tx[1:] != tx[:-1]
    run_start = np.maximum.accumulate(np.where(start, np.arange(n), 0))
    first = start | ((np.arange(n) - run_start) % SUPER_TILES == 0)
    firsts = np.nonzero(first)[0]
    tile_super = (np.cumsum(first) - 1).astype(np.int32)
    return np.concatenate([firsts, [n]]).astype(np.int32), tile_super


def max_tx_chains(chains: ChainPlan, tx_count: int) -> int:
    """Largest number of reflection chains (all depths) of one transmitter."""
    off = np.asarray(chains.offset_host, dtype=np.int64)
    if off.size < MAX_DEPTH * tx_count + 1:
        return 0
    per_tx = off[MAX_DEPTH * np.arange(1, tx_count + 1)] - off[MAX_DEPTH * np.arange(tx_count)]
    return int(per_tx.max()) if per_tx.size else 0


def _spread_bits(v: np.ndarray) -> np.ndarray:
    """Interleave zeros into the low 21 bits of ``v`` (bit i -> bit 2i)."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    for shift, mask in ((16, 0x0000FFFF0000FFFF), (8, 0x00FF00FF00FF00FF), (4, 0x0F0F0F0F0F0F0F0F),
                        (2, 0x3333333333333333), (1, 0x5555555555555555)):
        v = (v | (v << np.uint64(shift))) & np.uint64(mask)
    return v


def locality_permutation(points: np.ndarray) -> np.ndarray:
    """Order of ``points`` [R,3] along a Z-order curve over the ranks of their
    x and y coordinates: a regular grid falls into 8 x 8 blocks per 64
    consecutive entries, so the receivers of one wavefront are neighbours and
    take the same branches (same visible chains and edges, similar traversals).
    Ties keep input order, so the permutation is deterministic."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if pts.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    _, xr = np.unique(pts[:, 0], return_inverse=True)
    _, yr = np.unique(pts[:, 1], return_inverse=True)
    key = _spread_bits(xr.reshape(-1)) | (_spread_bits(yr.reshape(-1)) << np.uint64(1))
    return np.argsort(key, kind="stable")


def cell_order_table(rx_pos: np.ndarray, offsets: np.ndarray, shared: bool) -> np.ndarray:
    """Evaluation order of every cell (see wc_solve.cell_order): each
    transmitter's range is permuted onto itself by ``locality_permutation`` of
    its receivers."""
    offsets = np.asarray(offsets, dtype=np.int64)
    total = int(offsets[-1])
    order = np.empty(total, dtype=np.int64)
    if shared:
        perm = locality_permutation(rx_pos)
        for t in range(offsets.shape[0] - 1):
            order[offsets[t]:offsets[t + 1]] = offsets[t] + perm
    else:
        pts = np.asarray(rx_pos, dtype=np.float64).reshape(-1, 3)
        for t in range(offsets.shape[0] - 1):
            a, b = int(offsets[t]), int(offsets[t + 1])
            order[a:b] = a + locality_permutation(pts[a:b])
    return order.astype(np.int32)


def solve_inputs(compiled: CompiledScene, dscene: DeviceScene, *, components: int, max_depth: int,
                 coherent: bool, frequency_hz: float, tx_pos: np.ndarray | torch.Tensor | None = None,
                 tx_pol: np.ndarray | None = None, tx_power: np.ndarray | None = None,
                 rx_pos: np.ndarray | None = None, rx_pol: np.ndarray | None = None,
                 rx_count: np.ndarray | None = None, ray_count: int = DISCOVERY_RAY_COUNT,
                 chains: ChainPlan | None = None, lanes_per_cell: int = LANES_PER_CELL,
                 spatial_order: bool = True, visibility_mode: int = 0, cones: bool = True) -> SolveInputs:
    """Upload endpoint tables. Without ``rx_count`` every transmitter shares
    all receivers; with it, receivers are the concatenation of per-transmitter
    sets of the given sizes (uneven batched solves). ``lanes_per_cell`` lanes
    split each cell's reflection chains and diffraction edges, so a tile holds
    64 / lanes_per_cell cells. ``spatial_order`` evaluates cells in the
    Z-order of their receivers (outputs are unchanged, bit for bit).
    ``cones`` lets the chain culls test each chain's receiver cone first
    (outputs unchanged, bit for bit: the cone only drops chains the exact
    test rejects)."""
    if lanes_per_cell not in (1, 2, 4, 8, 16, 32, 64):
        raise ValueError("lanes_per_cell must be a power of two in [1, 64]")
    device = dscene.device
    ep = compiled.endpoints
    tx_pos = ep.tx_position if tx_pos is None else tx_pos
    if isinstance(tx_pos, torch.Tensor):
        tx_pos_t = tx_pos.detach().to(device=device, dtype=torch.float32).contiguous()
    else:
        tx_pos_t = torch.as_tensor(np.ascontiguousarray(tx_pos, dtype=np.float32)).to(device)
    t = int(tx_pos_t.shape[0])
    tx_pol = ep.tx_polarization if tx_pol is None else tx_pol
    tx_power = ep.tx_power if tx_power is None else tx_power
    rx_pos = ep.rx_position if rx_pos is None else rx_pos
    rx_pol = ep.rx_polarization if rx_pol is None else rx_pol
    f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).to(device)  # noqa: E731
    if rx_count is None:
        r = int(rx_pos.shape[0])
        offsets = np.arange(t + 1, dtype=np.int64) * r
        shared = True
        counts = np.full(t, r, dtype=np.int64)
    else:
        counts = np.asarray(rx_count, dtype=np.int64)
        offsets = np.concatenate([[0], np.cumsum(counts)])
        shared = False
    if offsets[-1] >= 2**31:
        raise RuntimeError("transmitter x receiver cell count exceeds int32 indexing")
    if chains is None:
        chains = plan_chains(compiled, dscene, tx_pos_t, max_depth, reflection=bool(components & 2),
                             ray_count=ray_count)
    cell_offset = torch.as_tensor(offsets.astype(np.int32)).to(device)
    rx_host = rx_pos.detach().cpu().numpy() if isinstance(rx_pos, torch.Tensor) else rx_pos
    order = (torch.as_tensor(cell_order_table(rx_host, offsets, shared)).to(device)
             if spatial_order and offsets[-1] > 0 else None)
    tiles_np = make_tiles(offsets, TILE // lanes_per_cell)
    tiles = torch.as_tensor(tiles_np if tiles_np.size else np.zeros((1, 2), np.int32)).to(device)
    # supertiles of the chain cull hold 16 tiles of 64 positions (one lane per cell)
    super_first, tile_super = super_tables(tiles_np if lanes_per_cell == 1 else np.zeros((0, 2), np.int32))
    super_first_t = torch.as_tensor(super_first).to(device)
    tile_super_t = torch.as_tensor(tile_super if tile_super.size else np.zeros(1, np.int32)).to(device)
    if lanes_per_cell != 1:
        super_first_t = None
    inputs = SolveInputs(tx_pos_t, f32(tx_pol).reshape(-1, 3), f32(tx_power).reshape(-1), f32(rx_pos).reshape(-1, 3),
                         f32(rx_pol).reshape(-1, 3), cell_offset, shared, chains, None, tiles,  # type: ignore[arg-type]
                         int(tiles_np.shape[0]), counts, order)
    p = _native.ptr
    inputs.view = _native.WcSolve(
        tx_pos=p(inputs.tx_pos), tx_pol=p(inputs.tx_pol), tx_power=p(inputs.tx_power), tx_count=t,
        rx_pos=p(inputs.rx_pos), rx_pol=p(inputs.rx_pol), rx_pol_stride=1, cell_offset=p(cell_offset),
        rx_shared=1 if shared else 0, total_cells=int(offsets[-1]), chain_offset=p(chains.offset),
        chain_groups=p(chains.groups), chain_images=p(chains.images), frequency_hz=float(frequency_hz),
        components=int(components), max_depth=int(max_depth), coherent=1 if coherent else 0,
        lanes_per_cell=int(lanes_per_cell), cell_order=p(order), max_tx_chains=max_tx_chains(chains, t),
        super_first_tile=p(super_first_t), tile_super=p(tile_super_t) if super_first_t is not None else None,
        super_count=int(super_first.shape[0] - 1) if super_first_t is not None else 0,
        visibility_mode=int(visibility_mode),
        chain_cones=p(chain_cones(dscene, chains)) if cones and components & 2 else None,
    )
    inputs.__dict__["_super_tables"] = (super_first_t, tile_super_t)
    return inputs


@dataclass
class RadiomapOutputs:
    power_total: torch.Tensor        # [cells] f32
    field_total: torch.Tensor        # [cells] c64
    component_power: torch.Tensor    # [4, cells] f32
    component_field: torch.Tensor    # [4, cells] c64
    diffraction_vector: torch.Tensor  # [cells, 3] c64
    path_count: torch.Tensor         # [cells] i32
    topology: torch.Tensor | None
    topology_capacity: int
    failure: torch.Tensor
    workspace: torch.Tensor          # per-cell partial sums between the two radiomap passes


def workspace_bytes(dscene: DeviceScene, inputs: SolveInputs) -> int:
    """Bytes of the radiomap workspace of one solve (wc_radiomap_workspace_bytes)."""
    return max(int(_native.call("wc_radiomap_workspace_bytes", inputs.total_cells, int(inputs.view.lanes_per_cell),
                                int(dscene.view.edge_count), int(inputs.view.max_tx_chains),
                                int(inputs.view.tx_count))), 32)


def radiomap_forward(dscene: DeviceScene, inputs: SolveInputs, *, topology_capacity: int = 0,
                     tiled: bool = True, out: RadiomapOutputs | None = None,
                     workspace: torch.Tensor | None = None) -> RadiomapOutputs:
    """Fused radiomap forward. ``workspace`` (uint8, at least
    ``workspace_bytes``) lets solves that run one after another share one
    buffer; by default each output set owns its own."""
    device = dscene.device
    cells = inputs.total_cells
    width = int(inputs.view.lanes_per_cell)
    if width > 1:
        tiled = True  # lane groups exist only in the tiled launch
    if topology_capacity and topology_capacity % width:
        topology_capacity += width - topology_capacity % width
    if out is None:
        out = RadiomapOutputs(
            power_total=torch.empty(cells, dtype=torch.float32, device=device),
            field_total=torch.empty(cells, dtype=torch.complex64, device=device),
            component_power=torch.empty((4, cells), dtype=torch.float32, device=device),
            component_field=torch.empty((4, cells), dtype=torch.complex64, device=device),
            diffraction_vector=torch.empty((cells, 3), dtype=torch.complex64, device=device),
            path_count=torch.empty(cells, dtype=torch.int32, device=device),
            topology=(torch.empty((cells, topology_capacity), dtype=torch.int32, device=device)
                      if topology_capacity > 0 else None),
            topology_capacity=topology_capacity,
            failure=torch.zeros(1, dtype=torch.int32, device=device),
            workspace=(workspace if workspace is not None else
                       torch.empty(workspace_bytes(dscene, inputs), dtype=torch.uint8, device=device)),
        )
        if out.workspace.numel() < workspace_bytes(dscene, inputs):
            raise ValueError("radiomap workspace is smaller than wc_radiomap_workspace_bytes")
    p = _native.ptr
    view = _out_view(out)
    counter = _tile_counter(inputs, device)
    _native.call("wc_radiomap_forward", dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(view),
                 p(inputs.tiles) if tiled else None, inputs.tile_count if tiled else 0,
                 p(counter) if tiled else None, _native.stream_handle(device))
    return out


MAX_COUPLED_CANDIDATES = 1_000_000  # reference interactions/coupled.py:47
COUPLED_KINDS = 7  # 1 R->D, 2 D->R, 4 D->D


@dataclass
class CoupledOutputs:
    field: torch.Tensor   # [cells] c64: sum of the coupled path fields (sqrt(P) x coefficient)
    power: torch.Tensor   # [cells]: sum of the coupled path gains
    count: torch.Tensor   # [cells] i32: coupled paths found
    count_dd: torch.Tensor  # [cells] i32: double-diffraction paths among them


def coupled_candidates_per_cell(dscene: DeviceScene, kinds: int = COUPLED_KINDS) -> int:
    return int(_native.call("wc_coupled_candidates_per_cell", int(dscene.group_count), int(dscene.edge_count),
                            int(kinds)))


def coupled_fields(dscene: DeviceScene, inputs: SolveInputs, *, candidate_limit: int = MAX_COUPLED_CANDIDATES,
                   kinds: int = COUPLED_KINDS) -> CoupledOutputs:
    """Coupled reflection-diffraction paths of every cell (components 3, 4
    and 7; reference interactions/coupled.py:93-230): the cells are streamed
    in blocks of at most ``candidate_limit`` (capped at 1,000,000)
    candidates, as the reference streams receiver blocks; a single cell whose
    candidates exceed the budget raises the reference's error."""
    device = dscene.device
    cells = inputs.total_cells
    per_cell = coupled_candidates_per_cell(dscene, kinds)
    limit = min(int(candidate_limit), MAX_COUPLED_CANDIDATES)
    if per_cell > limit:
        raise RuntimeError(f"coupled reflection-diffraction topology requires {per_cell} candidates, exceeding "
                           f"coupled_candidate_limit={limit}")
    out = CoupledOutputs(field=torch.zeros(cells, dtype=torch.complex64, device=device),
                         power=torch.zeros(cells, dtype=torch.float32, device=device),
                         count=torch.zeros(cells, dtype=torch.int32, device=device),
                         count_dd=torch.zeros(cells, dtype=torch.int32, device=device))
    if per_cell == 0 or cells == 0:
        return out
    block = max(1, limit // per_cell)
    ws_bytes = int(_native.call("wc_coupled_workspace_bytes", min(block, cells), int(dscene.group_count),
                                int(dscene.edge_count), int(kinds)))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    p = _native.ptr
    view = _native.WcCoupledOut(field=p(torch.view_as_real(out.field)), power=p(out.power), count=p(out.count),
                                count_dd=p(out.count_dd))
    stream = _native.stream_handle(device)
    for begin in range(0, cells, block):
        end = min(cells, begin + block)
        _native.call("wc_coupled_paths", dscene.ref(), ctypes.byref(inputs.view), begin, end, int(kinds),
                     ctypes.byref(view), p(ws), ws_bytes, stream)
    return out


def _out_view(out: RadiomapOutputs) -> _native.WcRadiomapOut:
    p = _native.ptr
    return _native.WcRadiomapOut(
        power_total=p(out.power_total), field_total=p(out.field_total), component_power=p(out.component_power),
        component_field=p(out.component_field), diffraction_vector=p(out.diffraction_vector),
        path_count=p(out.path_count), topology=p(out.topology), topology_capacity=out.topology_capacity,
        failure=p(out.failure), workspace=p(out.workspace),
    )


def _tile_counter(inputs: SolveInputs, device: torch.device) -> torch.Tensor:
    counter = inputs.__dict__.get("_tile_counter")
    if counter is None:
        counter = torch.zeros(2, dtype=torch.int32, device=device)
        inputs.__dict__["_tile_counter"] = counter
    return counter


def radiomap_pass(dscene: DeviceScene, inputs: SolveInputs, out: RadiomapOutputs, pass_index: int) -> None:
    """One kernel of the tiled forward: 1 chain validity, 2 LoS / reflection /
    transmission fields, 3 edge visibility, 4 diffraction fields (run in
    that order on one solve and ``out``)."""
    p = _native.ptr
    view = _out_view(out)
    counter = _tile_counter(inputs, dscene.device)
    _native.call("wc_radiomap_pass", dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(view), p(inputs.tiles),
                 inputs.tile_count, p(counter), int(pass_index), _native.stream_handle(dscene.device))


def forward_with_topology(dscene: DeviceScene, inputs: SolveInputs, *, capacity: int = 64,
                          tiled: bool = True, workspace: torch.Tensor | None = None) -> RadiomapOutputs:
    """Forward pass that stores every cell's path descriptors. When a cell
    holds more paths than ``capacity`` (failure bit 2) the descriptor buffer is
    regrown to the next power of two above the largest cell count and the
    forward is rerun, so the adjoint always sees the complete topology."""
    out = radiomap_forward(dscene, inputs, topology_capacity=capacity, tiled=tiled, workspace=workspace)
    if int(out.failure.item()) & 2:
        worst = int(out.path_count.max().item())
        # a single lane of a cell's lane group may hold all of its paths
        grown = int(inputs.view.lanes_per_cell) * (1 << max(worst - 1, 1).bit_length())
        out = radiomap_forward(dscene, inputs, topology_capacity=grown, tiled=tiled, workspace=workspace)
        if int(out.failure.item()) & 2:
            raise RuntimeError("path descriptor capacity still exceeded after regrowth")
    return out


def public_edge_ids(dscene: DeviceScene, rows: dict[str, torch.Tensor]) -> None:
    """Rewrite the edge ids of diffraction rows (``edge_id``, ``primitive_id``
    and the first sequence slot) from positions in the device edge table to
    the public edge ids of the compiled scene. The table is ordered by public
    id, so the canonical row order is unchanged."""
    if dscene.edge_count == 0 or rows["edge_id"].numel() == 0:
        return
    table = dscene.edge_public_id
    diff = rows["component_id"] == 2
    for name in ("edge_id", "primitive_id"):
        col = rows[name]
        rows[name] = torch.where(diff & (col >= 0), table[col.clamp(min=0).long()], col)
    seq = rows["primitive_sequence"]
    if seq.numel():
        first = seq[:, 0]
        seq[:, 0] = torch.where(diff & (first >= 0), table[first.clamp(min=0).long()], first)


ROW_FIELDS_INT = ("tx_id", "rx_id", "depth", "component_id", "primitive_id", "edge_id", "material_id")
ROW_FIELDS_FLOAT = ("path_length_m", "delay_s", "path_gain")


def pair_order(dscene: DeviceScene, inputs: SolveInputs) -> torch.Tensor:
    """Cell at each canonical pair position (receiver-major, transmitter-minor),
    sorted on the device once per solve inputs (it depends on the layout only)."""
    if inputs.pair_cell is None:
        cells = inputs.total_cells
        device = dscene.device
        pair_cell = torch.empty(max(cells, 1), dtype=torch.int32, device=device)
        if cells:
            ws = torch.empty(int(_native.call("wc_paths_compact_workspace_bytes", cells)), dtype=torch.uint8,
                             device=device)
            max_rx = int(np.max(inputs.rx_count)) if len(inputs.rx_count) else 0
            _native.call("wc_paths_pair_order", _native.ptr(inputs.cell_offset), int(inputs.view.tx_count), cells,
                         max_rx, _native.ptr(pair_cell), _native.ptr(ws), _native.stream_handle(device))
        inputs.pair_cell = pair_cell[:cells]
    return inputs.pair_cell


def export_rows(dscene: DeviceScene, inputs: SolveInputs, *, width: int,
                max_paths_per_pair: int | None = None) -> dict[str, Any]:
    """Explicit path rows in canonical order: pairs rx-major, tx-minor; inside a
    pair (depth, component, primitive sequence, primitive, edge), identity
    duplicates dropped, at most ``max_paths_per_pair`` rows per pair. Counting,
    pair offsets, the in-pair order and the compaction all run on the device;
    the host reads two totals (rows emitted, rows kept) to size the tables.
    ``pair_offset`` [cells + 1] (int64) and ``pair_cell`` [cells] (cell of each
    canonical pair) describe the kept rows' segmentation; ``dropped`` counts
    identity winners removed by the cap."""

    device = dscene.device
    cells = inputs.total_cells
    stream = _native.stream_handle(device)
    p = _native.ptr
    counts = torch.empty(max(cells, 1), dtype=torch.int32, device=device)
    _native.call("wc_paths_count", dscene.ref(), ctypes.byref(inputs.view), p(counts), stream)
    pair_cell = pair_order(dscene, inputs)
    ws = torch.empty(int(_native.call("wc_paths_compact_workspace_bytes", max(cells, 1))), dtype=torch.uint8,
                     device=device)
    row_offset = torch.empty(max(cells, 1), dtype=torch.int32, device=device)
    emitted_offset = torch.empty(cells + 1, dtype=torch.int32, device=device)
    totals = torch.zeros(2, dtype=torch.int64, device=device)
    _native.call("wc_paths_row_offsets", p(counts), p(pair_cell), cells, p(row_offset), p(emitted_offset),
                 p(totals), p(ws), stream)
    total = int(totals[0].item())
    if total >= 2 ** 31:
        raise RuntimeError(f"path export: {total} rows exceed the int32 row index")
    w = max(int(width), 1)
    k = max(total, 1)
    rows: dict[str, torch.Tensor] = {}
    for name in ROW_FIELDS_INT:
        rows[name] = torch.empty(k, dtype=torch.int32, device=device)
    for name in ROW_FIELDS_FLOAT:
        rows[name] = torch.empty(k, dtype=torch.float32, device=device)
    rows["primitive_sequence"] = torch.empty((k, w), dtype=torch.int32, device=device)
    rows["material_sequence"] = torch.empty((k, w), dtype=torch.int32, device=device)
    rows["interaction_positions"] = torch.empty((k, w, 3), dtype=torch.float32, device=device)
    rows["interaction_normals"] = torch.empty((k, w, 3), dtype=torch.float32, device=device)
    rows["path_field"] = torch.empty(k, dtype=torch.complex64, device=device)
    rows["coefficient"] = torch.empty(k, dtype=torch.complex64, device=device)
    rows["field_xyz"] = torch.empty((k, 3), dtype=torch.complex64, device=device)
    rows["arrival_direction"] = torch.empty((k, 3), dtype=torch.float32, device=device)
    rows["departure_direction"] = torch.empty((k, 3), dtype=torch.float32, device=device)
    view = _native.WcPathRows(width=w, **{name: p(rows[name]) for name in (
        *ROW_FIELDS_INT, *ROW_FIELDS_FLOAT, "primitive_sequence", "material_sequence", "interaction_positions",
        "interaction_normals", "path_field", "coefficient", "field_xyz", "arrival_direction",
        "departure_direction")})
    if total:
        _native.call("wc_paths_emit", dscene.ref(), ctypes.byref(inputs.view), p(row_offset), p(counts),
                     ctypes.byref(view), stream)
    perm = torch.empty(k, dtype=torch.int32, device=device)
    keep = torch.empty(k, dtype=torch.uint8, device=device)
    kept_offset = torch.empty(cells + 1, dtype=torch.int32, device=device)
    source_row = torch.empty(k, dtype=torch.int32, device=device)
    dropped = torch.zeros(1, dtype=torch.int32, device=device)
    cap = -1 if max_paths_per_pair is None else int(max_paths_per_pair)
    _native.call("wc_paths_compact", ctypes.byref(view), p(emitted_offset), cells, cap, p(perm), p(keep),
                 p(kept_offset), p(source_row), p(totals[1:]), p(dropped), p(ws), stream)
    kept_total, dropped_count = (int(x) for x in torch.cat([totals[1:], dropped.to(torch.int64)]).tolist())
    index = source_row[:kept_total].long()
    out: dict[str, Any] = {name: value.index_select(0, index) for name, value in rows.items()}
    public_edge_ids(dscene, out)
    out["pair_offset"] = kept_offset.to(torch.int64)
    out["pair_cell"] = pair_cell.to(torch.int64)
    out["dropped"] = dropped_count
    return out
