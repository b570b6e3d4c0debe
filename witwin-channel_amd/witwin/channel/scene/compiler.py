"""Scene compile: geometry, coplanar reflection groups, wedge edges, materials.

Produces a :class:`CompiledScene` whose host arrays (numpy) describe the
triangle soup, the coplanar face groups used by the image method, the
diffraction edge table and the material ABI (first-layer scalar view plus CSR
layer stacks), and whose :class:`DeviceScene` holds the same tables in HBM
together with the LBVH built by ``wc_bvh_build``.

Reference: witwin/channel/scene/compiler.py:131-322 (stores), :967-1150
(material records), native/channel/kernels/paths.cu:3142-3300 (coplanar face
groups keyed by surface id and the plane quantized at 1e-4),
witwin/channel/scene/resources.py:169-245 (scene + BVH creation).
"""

from __future__ import annotations

import ctypes
import dataclasses
import math
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from witwin.channel.constants import EPS0
from witwin.core import AntennaState, PhysicalMaterial, ReceiverGrid, Scene

DIELECTRIC_MODEL_ID = 1
PEC_MODEL_ID = 2
PEC_EFFECTIVE_SIGMA_E = 1.0e9
PLANE_GROUP_QUANTIZATION = 1.0e-4
# vertex positions are merged at this resolution when building the edge table
EDGE_WELD_QUANTIZATION = 1.0e-5
# two faces whose interior angle is within this of pi are coplanar: no wedge
COPLANAR_ANGLE_TOLERANCE = 1.0e-4


@dataclass
class MaterialTable:
    eps_r: np.ndarray
    sigma_e: np.ndarray
    mu_r: np.ndarray
    gain: np.ndarray
    thickness_m: np.ndarray
    model_id: np.ndarray
    layer_offset: np.ndarray
    layer_count: np.ndarray
    layer_thickness_m: np.ndarray
    layer_eps_r: np.ndarray
    layer_sigma_e: np.ndarray
    layer_mu_r: np.ndarray

    @property
    def count(self) -> int:
        return int(self.eps_r.shape[0])


@dataclass
class Endpoints:
    tx_position: np.ndarray      # [T,3] f32
    tx_polarization: np.ndarray  # [T,3] f32
    tx_power: np.ndarray         # [T] f32
    rx_position: np.ndarray      # [R,3] f32
    rx_polarization: np.ndarray  # [R,3] f32
    layout_kind: str             # "grid" (single receiver grid) or "point"
    grid_shape: tuple[int, int] | None


@dataclass
class CompiledScene:
    vertices: np.ndarray
    faces: np.ndarray
    face_material: np.ndarray
    face_surface: np.ndarray
    face_group: np.ndarray
    group_plane: np.ndarray
    group_face_offset: np.ndarray
    group_faces: np.ndarray
    group_material: np.ndarray
    representative_faces: np.ndarray
    materials: MaterialTable
    edges: np.ndarray            # structured array with the wc_edge layout
    frequency_hz: float
    endpoints: Endpoints
    edge_policy: str = "all"
    # [E, 4] int32 vertex ids (a, b, c0, c1) of every edge record (see edge_geometry)
    edge_vertices: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    # [E] int32 public id of every edge record: the id path rows report (record_ids)
    edge_public_id: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    # [F] structure of every face and [S] first vertex row of every structure
    face_structure: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    structure_vertex_base: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int64))
    device: "DeviceScene | None" = None
    cache: dict[str, Any] = field(default_factory=dict)

    @property
    def face_count(self) -> int:
        return int(self.faces.shape[0])

    @property
    def group_count(self) -> int:
        return int(self.group_plane.shape[0])

    @property
    def edge_count(self) -> int:
        return int(self.edges.shape[0])

    def face_vertices(self) -> np.ndarray:
        """[F, 9] float32: the three vertices of each face, original order."""
        return self.vertices[self.faces].reshape(-1, 9).astype(np.float32)


EDGE_DTYPE = np.dtype(
    [(name, np.float32) for name in (
        "ax", "ay", "az", "dx", "dy", "dz", "length", "n", "t0x", "t0y", "t0z", "n0x", "n0y", "n0z")]
    + [(name, np.int32) for name in ("face0", "face1", "group0", "group1", "material0", "material1")]
)


# ---------------------------------------------------------------- materials

def _sigma_from(eps_r: Any, sigma_e: float, frequency_hz: float) -> tuple[float, float]:
    """Real permittivity and total conductivity; a complex eps_r carries its
    loss as sigma = -Im(eps_r) * omega * eps0 (reference compiler.py:955-965)."""
    plain = lambda x: float(x.detach()) if isinstance(x, torch.Tensor) else float(x)  # noqa: E731
    if isinstance(eps_r, complex):
        extra = -eps_r.imag * (2.0 * math.pi * frequency_hz * EPS0)
        return float(eps_r.real), plain(sigma_e) + extra
    return plain(eps_r), plain(sigma_e)


def compile_materials(materials: list[PhysicalMaterial], frequency_hz: float) -> MaterialTable:
    rows = []
    layers = []
    offsets = []
    counts = []
    for material in materials:
        model = PEC_MODEL_ID if material.conductor_model == "perfect" else DIELECTRIC_MODEL_ID
        stack = material.layer_rows()
        offsets.append(len(layers))
        counts.append(len(stack))
        for layer in stack:
            eps, sigma = _sigma_from(layer.eps_r, layer.sigma_e, frequency_hz)
            if model == PEC_MODEL_ID:
                sigma = max(sigma, PEC_EFFECTIVE_SIGMA_E)
            layers.append((float(layer.thickness_m), eps, sigma, float(layer.mu_r)))
        first = layers[offsets[-1]]
        rows.append((first[1], first[2], first[3], float(material.gain), first[0], model))
    if not rows:
        rows.append((1.0, 0.0, 1.0, 1.0, 0.0, DIELECTRIC_MODEL_ID))
        offsets.append(0)
        counts.append(1)
        layers.append((0.0, 1.0, 0.0, 1.0))
    arr = np.asarray(rows, dtype=np.float64)
    lay = np.asarray(layers, dtype=np.float64)
    return MaterialTable(
        eps_r=arr[:, 0].astype(np.float32),
        sigma_e=arr[:, 1].astype(np.float32),
        mu_r=arr[:, 2].astype(np.float32),
        gain=arr[:, 3].astype(np.float32),
        thickness_m=arr[:, 4].astype(np.float32),
        model_id=arr[:, 5].astype(np.int32),
        layer_offset=np.asarray(offsets, dtype=np.int32),
        layer_count=np.asarray(counts, dtype=np.int32),
        layer_thickness_m=lay[:, 0].astype(np.float32),
        layer_eps_r=lay[:, 1].astype(np.float32),
        layer_sigma_e=lay[:, 2].astype(np.float32),
        layer_mu_r=lay[:, 3].astype(np.float32),
    )


# ---------------------------------------------------------------- coplanar groups

def coplanar_face_groups(tri_a: np.ndarray, normals: np.ndarray, surface_ids: np.ndarray,
                         quantization: float = PLANE_GROUP_QUANTIZATION) -> dict[str, Any]:
    """Group faces by (surface id, quantized canonical plane).

    The normal is flipped so its largest-magnitude component is positive; the
    key is (surface_id, round(n*s), round(offset*s)) with s = 1/quantization,
    and group ids follow the lexicographic order of the keys (faces with equal
    keys keep ascending face order). Mirrors the reference key construction of
    native/channel/kernels/paths.cu:3142-3190 and its sort-based assignment.
    """
    tri_a = np.asarray(tri_a, dtype=np.float32)
    normals = np.asarray(normals, dtype=np.float32)
    count = tri_a.shape[0]
    if count == 0:
        return {"face_group_id": np.zeros(0, np.int32), "representative_faces": np.zeros(0, np.int64),
                "group_count": 0}
    nrm = np.sqrt((normals.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32))
    inv = np.where(nrm > 1.0e-6, np.float32(1.0) / np.maximum(nrm, np.float32(1e-30)), np.float32(0.0))
    n = normals * inv[:, None]
    major = np.argmax(np.abs(n), axis=1)
    sign = np.where(n[np.arange(count), major] < 0.0, np.float32(-1.0), np.float32(1.0))
    n = n * sign[:, None]
    offset = (n * tri_a).sum(axis=1, dtype=np.float32)
    scale = np.float32(1.0 / quantization)
    keys = np.stack([
        np.asarray(surface_ids, dtype=np.int64),
        np.rint(n[:, 0] * scale).astype(np.int64),
        np.rint(n[:, 1] * scale).astype(np.int64),
        np.rint(n[:, 2] * scale).astype(np.int64),
        np.rint(offset * scale).astype(np.int64),
    ], axis=1)
    order = np.lexsort(keys.T[::-1])  # stable, first column most significant
    sorted_keys = keys[order]
    starts = np.ones(count, dtype=bool)
    starts[1:] = np.any(sorted_keys[1:] != sorted_keys[:-1], axis=1)
    group_of_sorted = np.cumsum(starts) - 1
    face_group = np.empty(count, dtype=np.int32)
    face_group[order] = group_of_sorted.astype(np.int32)
    representative = order[starts].astype(np.int64)
    return {"face_group_id": face_group, "representative_faces": representative,
            "group_count": int(starts.sum())}


def _group_tables(vertices: np.ndarray, faces: np.ndarray, face_group: np.ndarray, representative: np.ndarray,
                  face_material: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    g_count = representative.shape[0]
    v64 = vertices.astype(np.float64)
    planes = np.zeros((g_count, 4), dtype=np.float64)
    for g, f in enumerate(representative):
        a, b, c = v64[faces[f]]
        n = np.cross(b - a, c - a)
        n = n / max(np.linalg.norm(n), 1e-300)
        k = int(np.argmax(np.abs(n)))
        if n[k] < 0.0:
            n = -n
        planes[g, :3] = n
        planes[g, 3] = float(np.dot(n, a))
    order = np.argsort(face_group, kind="stable")
    counts = np.bincount(face_group, minlength=g_count)
    offsets = np.zeros(g_count + 1, dtype=np.int32)
    offsets[1:] = np.cumsum(counts)
    group_material = face_material[representative].astype(np.int32) if g_count else np.zeros(0, np.int32)
    return planes, offsets, order.astype(np.int32), group_material


# ---------------------------------------------------------------- edges

def edge_geometry(vertices: np.ndarray, edge_vertices: np.ndarray) -> dict[str, np.ndarray]:
    """Float64 geometry of edges given by their vertex ids [E, 4] = (a, b, c0,
    c1): a, b the end points in face 0's winding order, c0 the third vertex of
    face 0, c1 the third vertex of face 1 (-1 for a half-plane).

    Returns the origin A = v[a], unit direction d, length, the face-0
    in-plane direction t0 (perpendicular to d, towards c0), the face-0 normal
    n0 (winding normal with its d component removed) and the wedge index
    n = 2 - alpha / pi, alpha = atan2(-n0.t1, t0.t1) in [0, 2 pi) (2 for a
    half-plane). The wedge is a function of these four vertices only, which is
    what the vertex adjoint differentiates.
    """
    v = np.asarray(vertices, dtype=np.float64)
    ev = np.asarray(edge_vertices, dtype=np.int64).reshape(-1, 4)
    A = v[ev[:, 0]]
    d = v[ev[:, 1]] - A
    length = np.sqrt((d * d).sum(axis=1))
    d = d / np.maximum(length, 1e-300)[:, None]

    def in_plane(third: np.ndarray) -> np.ndarray:
        w = third - A
        w = w - d * (w * d).sum(axis=1)[:, None]
        return w / np.maximum(np.sqrt((w * w).sum(axis=1)), 1e-300)[:, None]

    c0 = v[ev[:, 2]]
    t0 = in_plane(c0)
    n0 = np.cross(v[ev[:, 1]] - A, c0 - A)
    n0 = n0 / np.maximum(np.sqrt((n0 * n0).sum(axis=1)), 1e-300)[:, None]
    n0 = n0 - d * (n0 * d).sum(axis=1)[:, None]
    n0 = n0 / np.maximum(np.sqrt((n0 * n0).sum(axis=1)), 1e-300)[:, None]
    wedge = ev[:, 3] >= 0
    t1 = in_plane(v[np.where(wedge, ev[:, 3], ev[:, 2])])
    alpha = np.arctan2(-(n0 * t1).sum(axis=1), (t0 * t1).sum(axis=1))
    alpha = np.where(alpha < 0.0, alpha + 2.0 * math.pi, alpha)
    n = np.where(wedge, 2.0 - alpha / math.pi, 2.0)
    return {"A": A, "d": d, "length": length, "t0": t0, "n0": n0, "alpha": alpha, "n": n, "wedge": wedge}


def _edge_rows(geometry: dict[str, np.ndarray], ids: np.ndarray) -> np.ndarray:
    """wc_edge records from edge_geometry output and [E, 6] integer columns
    (face0, face1, group0, group1, material0, material1)."""
    rows = np.zeros(ids.shape[0], dtype=EDGE_DTYPE)
    for k, axis in enumerate("xyz"):
        rows["a" + axis] = geometry["A"][:, k]
        rows["d" + axis] = geometry["d"][:, k]
        rows["t0" + axis] = geometry["t0"][:, k]
        rows["n0" + axis] = geometry["n0"][:, k]
    rows["length"] = geometry["length"]
    rows["n"] = geometry["n"]
    for k, name in enumerate(("face0", "face1", "group0", "group1", "material0", "material1")):
        rows[name] = ids[:, k]
    return rows


def record_ids(faces: np.ndarray, face_structure: np.ndarray | None,
               structure_vertex_base: np.ndarray | None) -> dict[tuple[int, int, int], int]:
    """Public edge id of every (structure, local vertex a, local vertex b), a < b.

    The reference numbers its edge records structure by structure: each
    structure contributes one record per distinct vertex pair of its faces,
    in ascending (lower, higher) local vertex id order, after the records of
    the structures before it (reference scene/resources.py:79-148 exports the
    records; their ids are what path rows carry as edge_id, e.g. the wedge of
    tests/deterministic/test_diffraction_single_wedge.py:41-50 yields ids
    0, 1, 2, 5 with record 3 the second copy of the shared edge). Without
    structure tables every face belongs to structure 0.
    """
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    if face_structure is None:
        face_structure = np.zeros(faces.shape[0], dtype=np.int64)
        structure_vertex_base = np.zeros(1, dtype=np.int64)
    face_structure = np.asarray(face_structure, dtype=np.int64)
    base = np.asarray(structure_vertex_base, dtype=np.int64)
    ids: dict[tuple[int, int, int], int] = {}
    cursor = 0
    for s in range(base.shape[0]):
        local = faces[face_structure == s] - base[s]
        if local.size == 0:
            continue
        pairs = np.concatenate([local[:, [0, 1]], local[:, [1, 2]], local[:, [2, 0]]], axis=0)
        pairs = np.unique(np.sort(pairs, axis=1), axis=0)
        for r, (a, b) in enumerate(pairs.tolist()):
            ids[(s, a, b)] = cursor + r
        cursor += pairs.shape[0]
    return ids


def build_edge_table(vertices: np.ndarray, faces: np.ndarray, face_group: np.ndarray, face_material: np.ndarray,
                     policy: str = "all", face_structure: np.ndarray | None = None,
                     structure_vertex_base: np.ndarray | None = None,
                     return_ids: bool = False) -> tuple[np.ndarray, ...]:
    """Diffraction edge table and the vertex ids of every edge.

    Every triangle edge is keyed by its welded end points. An edge shared by
    two faces of different, non-coplanar planes is a wedge whose exterior is the
    side the lower-id face's winding normal points to; an edge of exactly one
    face is a half-plane (n = 2). Edges between coplanar faces and edges with
    more than two faces are not diffracting. A wedge whose faces belong to two
    structures merges the two boundary records of the reference into the one
    with the lower id (reference scene/resources.py:286-404). Edges are ordered
    by their public ids (``record_ids``; the lower id of a merged pair).
    ``policy="vertical_only"`` keeps edges with |dz|/length > 0.7. Returns
    (records, [E, 4] int32 vertex ids (a, b, c0, c1), see edge_geometry) and,
    with ``return_ids``, the [E] int32 public ids.
    """
    v = vertices.astype(np.float64)
    q = np.rint(v / EDGE_WELD_QUANTIZATION).astype(np.int64)
    public = record_ids(faces, face_structure, structure_vertex_base)
    fstruct = (np.zeros(faces.shape[0], np.int64) if face_structure is None
               else np.asarray(face_structure, dtype=np.int64))
    vbase = (np.zeros(1, np.int64) if structure_vertex_base is None
             else np.asarray(structure_vertex_base, dtype=np.int64))
    table: dict[tuple, list] = {}
    order: list[tuple] = []
    for f in range(faces.shape[0]):
        idx = faces[f]
        for k in range(3):
            i0, i1, i2 = int(idx[k]), int(idx[(k + 1) % 3]), int(idx[(k + 2) % 3])
            a, b = tuple(q[i0]), tuple(q[i1])
            key = (a, b) if a <= b else (b, a)
            if key not in table:
                table[key] = []
                order.append(key)
            table[key].append((f, i0, i1, i2))

    def entry_id(f: int, i0: int, i1: int) -> int:
        s = int(fstruct[f])
        la, lb = i0 - int(vbase[s]), i1 - int(vbase[s])
        return public[(s, min(la, lb), max(la, lb))]

    cand_v, cand_i, cand_p = [], [], []
    for key in order:
        entries = table[key]
        if len(entries) > 2:
            continue
        f0, i0, i1, i2 = entries[0]
        if len(entries) == 2:
            f1, j0, j1, j2 = entries[1]
            if face_group[f1] == face_group[f0]:
                continue
            cand_v.append((i0, i1, i2, j2))
            cand_i.append((f0, f1, int(face_group[f0]), int(face_group[f1]), int(face_material[f0]),
                           int(face_material[f1])))
            cand_p.append(min(entry_id(f0, i0, i1), entry_id(f1, j0, j1)))
        else:
            cand_v.append((i0, i1, i2, -1))
            cand_i.append((f0, -1, int(face_group[f0]), -1, int(face_material[f0]), -1))
            cand_p.append(entry_id(f0, i0, i1))
    ev = np.asarray(cand_v, dtype=np.int64).reshape(-1, 4)
    ids = np.asarray(cand_i, dtype=np.int64).reshape(-1, 6)
    pid = np.asarray(cand_p, dtype=np.int64).reshape(-1)
    g = edge_geometry(v, ev)
    keep = g["length"] >= 1e-9
    if policy == "vertical_only":
        keep &= np.abs(g["d"][:, 2]) > 0.7
    keep &= ~(g["wedge"] & ((np.abs(g["alpha"] - math.pi) < COPLANAR_ANGLE_TOLERANCE) | (g["alpha"] < 1e-6)))
    sel = np.nonzero(keep)[0]
    sel = sel[np.argsort(pid[sel], kind="stable")]
    g = {k: a[sel] for k, a in g.items()}
    out = (_edge_rows(g, ids[sel]), ev[sel].astype(np.int32))
    return out + (pid[sel].astype(np.int32),) if return_ids else out


def build_edges(vertices: np.ndarray, faces: np.ndarray, face_group: np.ndarray, face_material: np.ndarray,
                policy: str = "all") -> np.ndarray:
    """Diffraction edge records (see build_edge_table)."""
    return build_edge_table(vertices, faces, face_group, face_material, policy)[0]


# ---------------------------------------------------------------- endpoints

def _pol(endpoint: AntennaState) -> np.ndarray:
    if endpoint.polarization is None:
        return np.array([0.0, 0.0, 1.0], dtype=np.float32)
    return endpoint.polarization.detach().cpu().numpy().astype(np.float32)


def grid_points(grid: ReceiverGrid) -> np.ndarray:
    """p = origin + x_axis*row*s0 + y_axis*col*s1, flat index row*cols+col, float32
    arithmetic as the reference grid kernel (native/channel/kernels/los.cu:391-420)."""
    rows, cols = grid.shape
    o = grid.origin.numpy().astype(np.float32)
    xa = grid.x_axis.numpy().astype(np.float32)
    ya = grid.y_axis.numpy().astype(np.float32)
    r = np.arange(rows, dtype=np.float32)[:, None] * np.float32(grid.spacing[0])
    c = np.arange(cols, dtype=np.float32)[None, :] * np.float32(grid.spacing[1])
    pts = np.empty((rows, cols, 3), dtype=np.float32)
    for k in range(3):
        pts[..., k] = o[k] + xa[k] * r + ya[k] * c
    return pts.reshape(-1, 3)


def compile_endpoints(scene: Scene) -> Endpoints:
    txs = scene.transmitters
    rxs = scene.receivers
    tx_pos = np.array([t.position.detach().cpu().numpy().reshape(3) for t in txs], dtype=np.float32).reshape(-1, 3)
    tx_pol = np.array([_pol(t) for t in txs], dtype=np.float32).reshape(-1, 3)
    tx_pow = np.array([1.0 if t.power_w is None else float(t.power_w) for t in txs], dtype=np.float32)
    positions = []
    pols = []
    for rx in rxs:
        pts = grid_points(rx) if isinstance(rx, ReceiverGrid) else rx.position.detach().cpu().numpy().reshape(-1, 3)
        positions.append(pts.astype(np.float32))
        pols.append(np.broadcast_to(_pol(rx), pts.shape).astype(np.float32))
    rx_pos = np.concatenate(positions, axis=0) if positions else np.zeros((0, 3), np.float32)
    rx_pol = np.concatenate(pols, axis=0) if pols else np.zeros((0, 3), np.float32)
    if len(rxs) == 1 and isinstance(rxs[0], ReceiverGrid):
        kind, shape = "grid", rxs[0].shape
    else:
        kind, shape = "point", None
    return Endpoints(tx_pos, tx_pol, tx_pow, rx_pos, rx_pol, kind, shape)


# ---------------------------------------------------------------- compile

def compile(scene: Scene, *, reference_frequency_hz: Any) -> CompiledScene:  # noqa: A001 - reference name
    frequency = float(reference_frequency_hz.detach()) if isinstance(reference_frequency_hz, torch.Tensor) \
        else float(reference_frequency_hz)
    if not frequency > 0.0:
        raise ValueError("reference_frequency_hz must be positive")
    vert_chunks, face_chunks, mat_ids, surf_ids, vert_leaves, face_struct, vbase = [], [], [], [], [], [], []
    unique: list[PhysicalMaterial] = []
    lookup: dict[int, int] = {}
    base = 0
    for index, structure in enumerate(scene.structures):
        mesh = structure.geometry
        v = mesh.vertices.detach().cpu().numpy().astype(np.float32)
        f = mesh.faces.detach().cpu().numpy().astype(np.int64)
        mat = structure.material
        if id(mat) not in lookup:
            lookup[id(mat)] = len(unique)
            unique.append(mat)
        vert_chunks.append(v)
        vert_leaves.append((mesh.vertices, base))
        face_chunks.append(f + base)
        face_struct.append(np.full(f.shape[0], index, dtype=np.int64))
        vbase.append(base)
        mat_ids.append(np.full(f.shape[0], lookup[id(mat)], dtype=np.int32))
        sid = structure.surface_id if structure.surface_id is not None else index
        surf_ids.append(np.full(f.shape[0], int(sid), dtype=np.int64))
        base += v.shape[0]
    vertices = np.concatenate(vert_chunks, axis=0) if vert_chunks else np.zeros((0, 3), np.float32)
    faces = np.concatenate(face_chunks, axis=0).astype(np.int32) if face_chunks else np.zeros((0, 3), np.int32)
    face_material = np.concatenate(mat_ids) if mat_ids else np.zeros(0, np.int32)
    face_surface = np.concatenate(surf_ids) if surf_ids else np.zeros(0, np.int64)
    materials = compile_materials(unique, frequency)
    firsts = [mat.layer_rows()[0] for mat in unique]
    ad_leaves = {
        "eps_r": [l.eps_r if isinstance(l.eps_r, torch.Tensor) else None for l in firsts],
        "sigma_e": [l.sigma_e if isinstance(l.sigma_e, torch.Tensor) else None for l in firsts],
        "tx_position": [t.position for t in scene.transmitters],
        "frequency": reference_frequency_hz if isinstance(reference_frequency_hz, torch.Tensor) else None,
        # (structure vertex tensor, first row in the concatenated vertex table)
        "vertices": vert_leaves,
    }
    if faces.shape[0]:
        tri = vertices[faces]
        normals = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]).astype(np.float32)
        groups = coplanar_face_groups(tri[:, 0], normals, face_surface)
    else:
        groups = coplanar_face_groups(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    face_group = groups["face_group_id"]
    representative = groups["representative_faces"]
    planes, offsets, members, group_material = _group_tables(vertices, faces, face_group, representative,
                                                             face_material)
    policy = getattr(scene.metadata.get("imported_edge_policy"), "edge_selection_mode", "all") \
        if scene.metadata else "all"
    face_structure = np.concatenate(face_struct) if face_struct else np.zeros(0, np.int64)
    structure_vertex_base = np.asarray(vbase if vbase else [0], dtype=np.int64)
    edges, edge_vertices, edge_public = build_edge_table(
        vertices, faces, face_group, face_material, policy, face_structure=face_structure,
        structure_vertex_base=structure_vertex_base, return_ids=True)
    compiled = CompiledScene(
        vertices=vertices, faces=faces, face_material=face_material, face_surface=face_surface,
        face_group=face_group, group_plane=planes, group_face_offset=offsets, group_faces=members,
        group_material=group_material, representative_faces=representative, materials=materials,
        edges=edges, frequency_hz=frequency, endpoints=compile_endpoints(scene), edge_policy=policy,
        edge_vertices=edge_vertices, edge_public_id=edge_public, face_structure=face_structure,
        structure_vertex_base=structure_vertex_base,
    )
    compiled.cache["ad_leaves"] = ad_leaves
    return compiled


def with_vertices(compiled: CompiledScene, vertices: np.ndarray) -> CompiledScene:
    """The compiled scene with moved vertices and the same topology: face
    groups, group membership and the edge list are kept; each group plane is
    recomputed from its representative face and each edge record from its
    four vertices. This is the geometry model the vertex adjoint
    differentiates (fixed topology)."""
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(compiled.vertices.shape))
    planes, _, _, _ = _group_tables(v, compiled.faces, compiled.face_group, compiled.representative_faces,
                                    compiled.face_material)
    edges = compiled.edges
    if compiled.edge_count:
        ids = np.stack([edges[k] for k in ("face0", "face1", "group0", "group1", "material0", "material1")],
                       axis=1).astype(np.int64)
        edges = _edge_rows(edge_geometry(v, compiled.edge_vertices), ids)
    return dataclasses.replace(compiled, vertices=v, group_plane=planes, edges=edges, device=None,
                               cache={})


def vertex_csr(vertex_count: int, faces: np.ndarray, group_face: np.ndarray,
               edge_vertices: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per-vertex lists of the geometry-adjoint contribution triples that
    touch it (see wc_geometry_vertex_grad): triple g*3 + k is vertex k of the
    representative face of group g, triple 3 G + e*4 + k vertex k of edge e.
    Returns (offset [V + 1], slot) int32; a vertex's slots keep ascending
    order, so the per-vertex sums are deterministic."""
    g_count = int(np.asarray(group_face).shape[0])
    gv = np.asarray(faces, dtype=np.int64)[np.asarray(group_face, dtype=np.int64)].reshape(-1)
    ev = np.asarray(edge_vertices, dtype=np.int64).reshape(-1)
    keep = ev >= 0
    verts = np.concatenate([gv, ev[keep]])
    slots = np.concatenate([np.arange(3 * g_count, dtype=np.int64),
                            3 * g_count + np.arange(ev.shape[0], dtype=np.int64)[keep]])
    order = np.argsort(verts, kind="stable")
    counts = np.bincount(verts, minlength=vertex_count)
    offset = np.zeros(vertex_count + 1, dtype=np.int64)
    offset[1:] = np.cumsum(counts)
    return offset.astype(np.int32), slots[order].astype(np.int32)


# ---------------------------------------------------------------- device scene

BVH_BUILDERS = ("sah", "lbvh")
BINARY_NODE_BYTES = 64   # wc_bvh_build / SAH build layout
WIDE_NODE_BYTES = 64     # the 4-wide nodes the tracers walk (wc_bvh_widen)


def wide_node_levels(nodes: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Refit order of a 4-wide tree (raw 64-byte nodes): the node ids reachable
    from the root grouped by depth, deepest level first, and the [levels + 1]
    offsets of the groups (the order wc_scene_refit refits them in)."""
    words = np.ascontiguousarray(nodes).view(np.int32).reshape(-1, WIDE_NODE_BYTES // 4)
    counts = words[:, 3].view(np.uint32) >> 24
    levels: list[list[int]] = [[0]]
    while True:
        nxt = [int(c) for n in levels[-1] for c in words[n, 12:12 + int(counts[n])] if c >= 0]
        if not nxt:
            break
        levels.append(nxt)
    order = np.array([n for level in reversed(levels) for n in level], dtype=np.int32)
    offsets = np.cumsum([0] + [len(level) for level in reversed(levels)]).astype(np.int32)
    return order, offsets


def group_hulls(vertices: np.ndarray, faces: np.ndarray, group_face_offset: np.ndarray, group_faces: np.ndarray,
                group_plane: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Convex hull of every coplanar group in its plane: ([G + 1] int32 CSR
    offsets, [H, 3] float64 points on the plane, counter-clockwise about the
    normal). A one-triangle group is its triangle; larger groups take the
    monotone-chain hull of their vertices in an in-plane basis."""
    g = int(group_plane.shape[0])
    n = group_plane[:, :3].astype(np.float64)
    off = group_plane[:, 3].astype(np.float64)
    # in-plane basis: u = a unit vector orthogonal to n, w = n x u
    helper = np.where(np.abs(n[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    u = np.cross(n, helper)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.cross(n, u)
    origin = n * off[:, None]
    counts = np.diff(group_face_offset)
    hull_pts: list[np.ndarray] = [np.zeros((0, 3))] * g
    v64 = vertices.astype(np.float64)

    def to_plane(gi: int, pts: np.ndarray) -> np.ndarray:
        d = pts - origin[gi]
        return np.stack([d @ u[gi], d @ w[gi]], axis=1)

    def from_plane(gi: int, uv: np.ndarray) -> np.ndarray:
        return origin[gi] + uv[:, :1] * u[gi] + uv[:, 1:2] * w[gi]

    def monotone(uv: np.ndarray) -> np.ndarray:
        pts = np.unique(uv, axis=0)
        if pts.shape[0] < 3:
            return pts

        def half(seq):
            out: list[np.ndarray] = []
            for p in seq:
                while len(out) >= 2 and ((out[-1][0] - out[-2][0]) * (p[1] - out[-2][1])
                                         - (out[-1][1] - out[-2][1]) * (p[0] - out[-2][0])) <= 0.0:
                    out.pop()
                out.append(p)
            return out

        lower = half(pts)
        upper = half(pts[::-1])
        return np.array(lower[:-1] + upper[:-1])

    single = np.nonzero(counts == 1)[0]
    if single.size:
        tri = v64[faces[group_faces[group_face_offset[single]]]]  # [S, 3, 3]
        for k, gi in enumerate(single):
            uv = to_plane(gi, tri[k])
            area = (uv[1, 0] - uv[0, 0]) * (uv[2, 1] - uv[0, 1]) - (uv[1, 1] - uv[0, 1]) * (uv[2, 0] - uv[0, 0])
            if area < 0.0:
                uv = uv[::-1]
            hull_pts[gi] = from_plane(gi, uv)
    for gi in np.nonzero(counts > 1)[0]:
        members = group_faces[group_face_offset[gi]:group_face_offset[gi + 1]]
        pts = v64[faces[members].reshape(-1)]
        hull_pts[gi] = from_plane(gi, monotone(to_plane(gi, pts)))
    sizes = np.array([h.shape[0] for h in hull_pts], dtype=np.int64)
    offset = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    points = np.concatenate(hull_pts, axis=0) if g else np.zeros((0, 3))
    return offset, np.ascontiguousarray(points, dtype=np.float64)


class DeviceScene:
    """HBM copy of a compiled scene plus its BVH, and the wc_scene view.

    ``bvh="sah"`` (default) builds a binned surface-area-heuristic tree on the
    host (``wc_bvh_build_sah``): fewer nodes per segment walk for every
    visibility test. ``bvh="lbvh"`` builds the Morton-order tree on the
    device (``wc_bvh_build``), for geometry that is rebuilt every step. Either
    binary tree is collapsed on the host into the 4-wide nodes the tracers
    walk. Both trees answer every visibility and hit query identically; they
    differ only in how many nodes a walk visits."""

    def __init__(self, compiled: CompiledScene, device: torch.device, bvh: str = "sah") -> None:
        from witwin.channel import _native

        if bvh not in BVH_BUILDERS:
            raise ValueError(f"bvh must be one of {BVH_BUILDERS}, got {bvh!r}")
        self.device = device
        self.bvh = bvh
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a)).to(device=device, dtype=dt).contiguous()  # noqa: E731
        mats = compiled.materials
        self.vertices = t(compiled.vertices, torch.float32)
        self.faces = t(compiled.faces, torch.int32)
        self.face_group = t(compiled.face_group, torch.int32)
        self.face_material = t(compiled.face_material, torch.int32)
        self.group_plane = t(compiled.group_plane, torch.float64)
        self.group_face_offset = t(compiled.group_face_offset, torch.int32)
        self.group_faces = t(compiled.group_faces, torch.int32)
        self.group_material = t(compiled.group_material, torch.int32)
        self.mat_eps = t(mats.eps_r, torch.float32)
        self.mat_sigma = t(mats.sigma_e, torch.float32)
        self.mat_mu = t(mats.mu_r, torch.float32)
        self.mat_gain = t(mats.gain, torch.float32)
        self.mat_thickness = t(mats.thickness_m, torch.float32)
        self.layer_offset = t(mats.layer_offset, torch.int32)
        self.layer_count = t(mats.layer_count, torch.int32)
        self.layer_thickness = t(mats.layer_thickness_m, torch.float32)
        self.layer_eps = t(mats.layer_eps_r, torch.float32)
        self.layer_sigma = t(mats.layer_sigma_e, torch.float32)
        self.layer_mu = t(mats.layer_mu_r, torch.float32)
        self.group_face = t(compiled.representative_faces, torch.int32)
        self.edge_vertices = t(compiled.edge_vertices.reshape(-1, 4), torch.int32)
        public = compiled.edge_public_id if compiled.edge_public_id.shape[0] == compiled.edge_count \
            else np.arange(compiled.edge_count)
        self.edge_public_id = t(public.reshape(-1), torch.int32)
        self.group_count = compiled.group_count
        self.edge_count = compiled.edge_count
        self._csr: tuple[torch.Tensor, torch.Tensor] | None = None
        self._hulls: tuple[torch.Tensor, torch.Tensor] | None = None
        edge_bytes = compiled.edges.view(np.uint8).reshape(-1) if compiled.edge_count else np.zeros(0, np.uint8)
        self.edges = torch.as_tensor(edge_bytes.copy()).to(device)
        f = compiled.face_count
        self.node_count = int(_native.call("wc_bvh_node_count", f))
        self.nodes = torch.zeros(self.node_count * WIDE_NODE_BYTES, dtype=torch.uint8, device=device)
        self.sorted_tris = torch.zeros(max(f, 1) * 48, dtype=torch.uint8, device=device)
        self.face_tris = torch.zeros(max(f, 1) * 48, dtype=torch.uint8, device=device)
        if f and bvh == "sah":
            host = lambda a, dt: np.ascontiguousarray(a, dtype=dt)  # noqa: E731
            verts, tri_faces = host(compiled.vertices, np.float32), host(compiled.faces, np.int32)
            groups, materials = host(compiled.face_group, np.int32), host(compiled.face_material, np.int32)
            nodes = np.zeros(self.node_count * WIDE_NODE_BYTES, np.uint8)
            sorted_tris = np.zeros(f * 48, np.uint8)
            face_tris = np.zeros(f * 48, np.uint8)
            _native.call("wc_bvh_build_sah", verts.ctypes.data, tri_faces.ctypes.data, groups.ctypes.data,
                         materials.ctypes.data, f, nodes.ctypes.data, sorted_tris.ctypes.data, face_tris.ctypes.data)
            self.nodes.copy_(torch.from_numpy(nodes))
            self.sorted_tris.copy_(torch.from_numpy(sorted_tris))
            self.face_tris.copy_(torch.from_numpy(face_tris))
            wide_host = nodes
        elif f:
            ws_bytes = int(_native.call("wc_bvh_workspace_bytes", f))
            workspace = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
            binary = torch.zeros(self.node_count * BINARY_NODE_BYTES, dtype=torch.uint8, device=device)
            _native.call("wc_bvh_build", _native.ptr(self.vertices), _native.ptr(self.faces),
                         _native.ptr(self.face_group), _native.ptr(self.face_material), f,
                         _native.ptr(binary), _native.ptr(self.sorted_tris), _native.ptr(self.face_tris),
                         _native.ptr(workspace), ws_bytes, _native.stream_handle(device))
            binary_host = binary.cpu().numpy()  # synchronizes with the build
            del workspace, binary
            wide = np.zeros(self.node_count * WIDE_NODE_BYTES, np.uint8)
            _native.call("wc_bvh_widen", binary_host.ctypes.data, f, wide.ctypes.data, None)
            self.nodes.copy_(torch.from_numpy(wide))
            wide_host = wide
        self._refit_order: torch.Tensor | None = None
        self._refit_offsets = np.zeros(1, np.int32)
        if f:
            order, self._refit_offsets = wide_node_levels(wide_host)
            self._refit_order = torch.as_tensor(order).to(device)
        p = _native.ptr
        self.view = _native.WcScene(
            nodes=p(self.nodes), tris=p(self.sorted_tris), tri_count=f, node_count=self.node_count,
            faces=p(self.face_tris), face_group=p(self.face_group), face_material=p(self.face_material),
            face_count=f, group_plane=p(self.group_plane), group_face_offset=p(self.group_face_offset),
            group_faces=p(self.group_faces), group_material=p(self.group_material),
            group_count=compiled.group_count, mat_eps=p(self.mat_eps), mat_sigma=p(self.mat_sigma),
            mat_mu=p(self.mat_mu), mat_gain=p(self.mat_gain), mat_thickness=p(self.mat_thickness),
            mat_layer_offset=p(self.layer_offset), mat_layer_count=p(self.layer_count),
            layer_thickness=p(self.layer_thickness), layer_eps=p(self.layer_eps), layer_sigma=p(self.layer_sigma),
            layer_mu=p(self.layer_mu), material_count=mats.count, edges=p(self.edges),
            edge_count=compiled.edge_count,
        )

    def ref(self) -> "ctypes._Pointer":
        return ctypes.byref(self.view)

    def refit(self, vertices: torch.Tensor) -> None:
        """Move the vertices (same topology) on the device: triangle tables,
        the tree's boxes, group planes and edge records are recomputed in place
        by ``wc_scene_refit``; faces, groups, the edge list and the tree's
        shape stay. The host CompiledScene is not updated
        (``with_vertices`` gives its host counterpart)."""
        from witwin.channel import _native

        v = vertices.detach().to(device=self.device, dtype=torch.float32).reshape(self.vertices.shape)
        self.vertices.copy_(v)
        self._hulls = None  # group hulls follow the vertices
        failure = torch.zeros(1, dtype=torch.int32, device=self.device)
        p = _native.ptr
        levels = np.ascontiguousarray(self._refit_offsets, dtype=np.int32)
        refit = _native.WcRefit(
            vertices=p(self.vertices), faces=p(self.faces), group_face=p(self.group_face),
            edge_vertices=p(self.edge_vertices), node_order=p(self._refit_order) if self._refit_order is not None
            else None, level_offset_host=levels.ctypes.data, levels=len(levels) - 1, failure=p(failure))
        _native.call("wc_scene_refit", self.ref(), ctypes.byref(refit), _native.stream_handle(self.device))
        if int(failure.item()):
            raise RuntimeError("wc_scene_refit: a node box could not be quantized; rebuild the scene")

    def hulls(self) -> tuple[torch.Tensor, torch.Tensor]:
        """Device copy of group_hulls for the current vertices (built on first
        use, rebuilt after a refit)."""
        if self._hulls is None:
            offset, points = group_hulls(self.vertices.cpu().numpy(), self.faces.cpu().numpy(),
                                         self.group_face_offset.cpu().numpy(), self.group_faces.cpu().numpy(),
                                         self.group_plane.cpu().numpy())
            self._hulls = (torch.as_tensor(offset).to(self.device),
                           torch.as_tensor(points if points.size else np.zeros((1, 3))).to(self.device))
        return self._hulls

    def vertex_csr(self) -> tuple[torch.Tensor, torch.Tensor]:
        """Device copy of vertex_csr for this scene (built on first use)."""
        if self._csr is None:
            offset, slot = vertex_csr(int(self.vertices.shape[0]), self.faces.cpu().numpy(),
                                      self.group_face.cpu().numpy(), self.edge_vertices.cpu().numpy())
            self._csr = (torch.as_tensor(offset).to(self.device), torch.as_tensor(slot).to(self.device))
        return self._csr


def device_scene(compiled: CompiledScene, device: torch.device, bvh: str = "sah") -> DeviceScene:
    """The compiled scene's device copy (cached on the scene per device and
    BVH builder)."""
    if compiled.device is None or compiled.device.device != device or compiled.device.bvh != bvh:
        compiled.device = DeviceScene(compiled, device, bvh=bvh)
    return compiled.device
