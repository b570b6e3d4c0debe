"""Seeded synthetic workloads used by the parity tests and the benchmark.

* ``two_ray_scene`` — one isotropic transmitter at (0, 0, 10) m over a flat
  ground (eps_r 5.31, sigma 0.0326 S/m, 5 m slab so the slab coefficient equals
  the half-space one to < 1e-11), 3.5 GHz, 64 x 64 receiver grid at z = 1.5 m
  with 1 m spacing centred on the transmitter.
* ``urban_canyon_scene`` — seeded 8 x 8 Manhattan grid of box buildings
  (30 x 30 m footprint, 15 m streets, heights U[10, 40] m, 12 triangles each)
  on a ground quad, concrete walls (eps_r 5.31, sigma 0.0326) with a seeded
  25 % of wall faces glass (eps_r 6.27, sigma 0.0043), 0.2 m thick; one
  transmitter at a street intersection, 25 m high; a receiver grid at
  z = 1.5 m covering the block (256 x 256 at 2 m by default).
* ``batched_transmitters`` — seeded transmitter positions over the streets
  with log-uniform receiver-set sizes.
* ``wedge_diffraction_scene`` — two perfectly conducting triangles (two
  structures) meeting along the vertical edge x = 2, y = 0, z in [-1, 2],
  transmitter (0, -1, 0.5), receiver (3, 1, 0.5), used at 3 GHz (the geometry
  of the reference fixture tests/support/scenes.py:172-215).
* ``heightfield_scene`` — seeded fBm terrain (128 x 128 vertices at 2 m
  pitch, relief 20 m, 32,258 triangles) with 40 random boxes (480
  triangles), 4 materials (soil, concrete, glass, brick), one transmitter
  25 m above the highest terrain point and a 256 x 256 receiver grid 1.5 m
  above it; diffraction edges follow the vertical-only edge policy.
"""

from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np
import torch

from witwin.core import (
    AntennaState,
    MaterialLayer,
    Mesh,
    PhysicalMaterial,
    ReceiverGrid,
    Scene,
    Structure,
    new_antenna_id,
)

CONCRETE = dict(eps_r=5.31, sigma_e=0.0326)
GLASS = dict(eps_r=6.27, sigma_e=0.0043)

# box faces as quads (outward winding): -x, +x, -y, +y, bottom, top
_BOX_QUADS = (
    ((0, 4, 6, 2), "wall"),   # x = x0
    ((1, 3, 7, 5), "wall"),   # x = x1
    ((0, 1, 5, 4), "wall"),   # y = y0
    ((2, 6, 7, 3), "wall"),   # y = y1
    ((0, 2, 3, 1), "floor"),  # z = 0
    ((4, 5, 7, 6), "roof"),   # z = h
)


def box_mesh(x0: float, y0: float, x1: float, y1: float, h: float) -> tuple[np.ndarray, np.ndarray, list[str]]:
    """12-triangle box with outward-facing winding; returns vertices, faces and
    the role of each face pair."""
    v = np.array([[x, y, z] for z in (0.0, h) for y in (y0, y1) for x in (x0, x1)], dtype=np.float32)
    faces, roles = [], []
    for (a, b, c, d), role in _BOX_QUADS:
        faces.append((a, b, c))
        faces.append((a, c, d))
        roles.extend([role, role])
    return v, np.asarray(faces, dtype=np.int64), roles


def ground_structure(x0: float, y0: float, x1: float, y1: float, material: PhysicalMaterial,
                     surface_id: int = 0) -> Structure:
    v = np.array([[x0, y0, 0.0], [x1, y0, 0.0], [x1, y1, 0.0], [x0, y1, 0.0]], dtype=np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)
    return Structure(Mesh(torch.from_numpy(v), torch.from_numpy(f)), material, name="ground", surface_id=surface_id)


def transmitter(position, power_w: float = 1.0, polarization=None) -> AntennaState:
    return AntennaState(new_antenna_id(), "tx", torch.tensor(position, dtype=torch.float32),
                        polarization=polarization, power_w=power_w)


def receiver(position, polarization=None) -> AntennaState:
    return AntennaState(new_antenna_id(), "rx", torch.tensor(position, dtype=torch.float32), polarization=polarization)


def grid(origin, shape, spacing) -> ReceiverGrid:
    return ReceiverGrid(new_antenna_id(), torch.tensor(origin, dtype=torch.float32),
                        torch.tensor([1.0, 0.0, 0.0]), torch.tensor([0.0, 1.0, 0.0]), shape, spacing)


def two_ray_scene(shape: tuple[int, int] = (64, 64), spacing: float = 1.0, extent: float = 1.0e4) -> Scene:
    ground = PhysicalMaterial(eps_r=5.31, sigma_e=0.0326, thickness_m=5.0, name="ground")
    rows, cols = shape
    origin = (-(rows - 1) * spacing / 2.0, -(cols - 1) * spacing / 2.0, 1.5)
    return Scene(structures=[ground_structure(-extent, -extent, extent, extent, ground)],
                 endpoints=[transmitter((0.0, 0.0, 10.0)), grid(origin, shape, spacing)])


def manhattan_structures(seed: int = 2, blocks: int = 8, footprint: float = 30.0, street: float = 15.0,
                         height_range: tuple[float, float] = (10.0, 40.0), glass_fraction: float = 0.25,
                         thickness: float = 0.2) -> tuple[list[Structure], float]:
    """Box buildings on a regular street grid; returns (structures, pitch)."""
    rng = np.random.default_rng(seed)
    concrete = PhysicalMaterial(thickness_m=thickness, name="concrete", **CONCRETE)
    glass = PhysicalMaterial(thickness_m=thickness, name="glass", **GLASS)
    pitch = footprint + street
    structures = []
    sid = 1
    for i in range(blocks):
        for j in range(blocks):
            x0 = street + i * pitch
            y0 = street + j * pitch
            h = float(rng.uniform(*height_range))
            v, f, roles = box_mesh(x0, y0, x0 + footprint, y0 + footprint, h)
            glass_faces = [k for k in range(0, 8, 2) if rng.uniform() < glass_fraction]
            # split walls by material: one structure per material so each face keeps its surface
            concrete_faces, glass_faces_list = [], []
            for k in range(0, 12, 2):
                target = glass_faces_list if (roles[k] == "wall" and k in glass_faces) else concrete_faces
                target.extend([f[k], f[k + 1]])
            for faces, mat in ((concrete_faces, concrete), (glass_faces_list, glass)):
                if not faces:
                    continue
                structures.append(Structure(Mesh(torch.from_numpy(v), torch.from_numpy(np.asarray(faces))), mat,
                                            name=f"box-{i}-{j}", surface_id=sid))
                sid += 1
    return structures, pitch


# street intersection (x, y in [45, 60] with the 15 m first street) offset from
# its centre so no transmitter->box-corner line passes exactly through another
# corner of the regular block lattice
URBAN_CANYON_TX = (53.231, 52.013, 25.0)


def urban_canyon_scene(seed: int = 2, shape: tuple[int, int] = (256, 256), spacing: float = 2.0,
                       tx_position=URBAN_CANYON_TX) -> Scene:
    structures, pitch = manhattan_structures(seed)
    extent = 15.0 + 8 * pitch
    ground = PhysicalMaterial(thickness_m=0.2, name="ground-concrete", **CONCRETE)
    structures.insert(0, ground_structure(-5.0, -5.0, extent + 5.0, extent + 5.0, ground))
    # grid centred on the block; origin offset by a fraction of the spacing so receivers do not sit
    # exactly on wall planes
    rows, cols = shape
    span_r, span_c = (rows - 1) * spacing, (cols - 1) * spacing
    origin = ((extent - span_r) / 2.0 + 0.173, (extent - span_c) / 2.0 + 0.291, 1.5)
    return Scene(structures=structures, endpoints=[transmitter(tx_position), grid(origin, shape, spacing)])


def street_points(rng: np.random.Generator, count: int, street: float = 15.0, pitch: float = 45.0,
                  blocks: int = 8, z: float | tuple[float, float] = 1.5) -> np.ndarray:
    """Uniform points on the street network (outside every footprint)."""
    extent = street + blocks * pitch
    pts = []
    while len(pts) < count:
        x, y = rng.uniform(0.0, extent, size=2)
        in_x = (x % pitch) >= street
        in_y = (y % pitch) >= street
        if in_x and in_y and x < extent and y < extent:
            continue  # inside a building footprint
        zz = rng.uniform(*z) if isinstance(z, tuple) else z
        pts.append((x, y, zz))
    return np.asarray(pts, dtype=np.float32)


def batched_transmitters(seed: int = 3, tx_count: int = 64, min_rx: int = 16, max_rx: int = 4096):
    """Transmitters over the streets (z U[3, 30]) and log-uniform receiver-set
    sizes; returns (tx_positions [T,3], rx_counts [T], rx_positions [sum,3])."""
    rng = np.random.default_rng(seed)
    tx = street_points(rng, tx_count, z=(3.0, 30.0))
    sizes = np.floor(np.exp(rng.uniform(math.log(min_rx), math.log(max_rx), size=tx_count))).astype(np.int64)
    rx = np.concatenate([street_points(rng, int(n)) for n in sizes], axis=0)
    return tx, sizes, rx


def wedge_diffraction_scene(edge_policy: str | None = None, tx=(0.0, -1.0, 0.5), rx=(3.0, 1.0, 0.5)) -> Scene:
    """PEC wedge of two single-triangle structures sharing one vertical edge
    (geometry of reference tests/support/scenes.py:172-215). Public edge ids:
    0 the shared vertical wedge, 1 and 2 the horizontal and slanted outline
    edges of the first structure, 3 the second structure's copy of the shared
    edge (merged into 0), 4 and 5 its horizontal and slanted outline edges."""
    pec = PhysicalMaterial.perfect_conductor()
    face_a = Structure(Mesh(torch.tensor([[2.0, 0.0, -1.0], [2.0, 0.0, 2.0], [2.0, 2.0, -1.0]]),
                            torch.tensor([[0, 1, 2]])), pec, name="wedge-a", surface_id=2)
    face_b = Structure(Mesh(torch.tensor([[2.0, 0.0, -1.0], [2.0, 0.0, 2.0], [4.0, 0.0, -1.0]]),
                            torch.tensor([[0, 2, 1]])), pec, name="wedge-b", surface_id=3)
    metadata = {} if edge_policy is None else {
        "imported_edge_policy": SimpleNamespace(edge_selection_mode=edge_policy)}
    return Scene(structures=[face_a, face_b], endpoints=[transmitter(tx), receiver(rx)], metadata=metadata)


def coupled_wall_wedge_scene(material: PhysicalMaterial | None = None, tx=(0.0, -2.0, 1.0), rx=(0.0, 2.0, 5.0),
                             ground: bool = True, grid_shape: tuple[int, int] | None = None,
                             grid_origin=(-3.0, -3.0, 3.0), grid_spacing: float = 0.5) -> Scene:
    """A reflector at z = 0 ([-5, 5]^2) and an L-shaped wedge whose convex edge
    runs along (2, y, 2), y in [-1, 1]: a horizontal face at z = 2 over
    x in [2, 4] and a vertical face at x = 2 over z in [2, 4] (geometry of
    reference tests/support/scenes.py:218-264), PEC unless ``material`` is
    given. ``ground=False`` drops the reflector; ``grid_shape`` replaces the
    single receiver by a receiver grid."""
    mat = PhysicalMaterial.perfect_conductor() if material is None else material
    v = [[2.0, -1.0, 2.0], [2.0, 1.0, 2.0], [4.0, -1.0, 2.0], [4.0, 1.0, 2.0], [2.0, -1.0, 4.0], [2.0, 1.0, 4.0]]
    f = [[0, 2, 3], [0, 3, 1], [0, 1, 5], [0, 5, 4]]
    if ground:
        v = [[-5.0, -5.0, 0.0], [5.0, -5.0, 0.0], [5.0, 5.0, 0.0], [-5.0, 5.0, 0.0]] + v
        f = [[0, 1, 2], [0, 2, 3]] + [[a + 4, b + 4, c + 4] for a, b, c in f]
    wedge = Structure(Mesh(torch.tensor(v, dtype=torch.float32), torch.tensor(f)), mat, name="wall-wedge",
                      surface_id=0)
    rx_end = receiver(rx) if grid_shape is None else grid(grid_origin, grid_shape, grid_spacing)
    return Scene(structures=[wedge], endpoints=[transmitter(tx), rx_end])


SOIL = dict(eps_r=15.0, sigma_e=0.035)
BRICK = dict(eps_r=3.75, sigma_e=0.038)


def fbm_heights(seed: int, size: int = 128, octaves: int = 6, relief: float = 20.0) -> np.ndarray:
    """[size, size] fractional-Brownian-motion heights in [0, relief]: value
    noise (seeded uniform lattices, bilinear interpolation) summed over
    octaves of doubling frequency and halving weight."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, size)
    total = np.zeros((size, size))
    for o in range(octaves):
        cells = 2 ** (o + 1)
        lattice = rng.uniform(-1.0, 1.0, (cells + 1, cells + 1))
        x = u * cells
        i = np.minimum(np.floor(x).astype(np.int64), cells - 1)
        f = x - i
        row = lattice[i] * (1.0 - f)[:, None] + lattice[i + 1] * f[:, None]
        total += 0.5 ** o * (row[:, i] * (1.0 - f)[None, :] + row[:, i + 1] * f[None, :])
    return relief * (total - total.min()) / (total.max() - total.min())


def heightfield_mesh(heights: np.ndarray, pitch: float) -> tuple[np.ndarray, np.ndarray]:
    """Terrain vertices (x = i pitch, y = j pitch, z = heights[i, j]) and two
    upward-facing triangles per lattice square."""
    n = heights.shape[0]
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v = np.stack([ii * pitch, jj * pitch, heights], axis=-1).reshape(-1, 3).astype(np.float32)
    a = (ii[:-1, :-1] * n + jj[:-1, :-1]).reshape(-1)
    b = a + n          # (i + 1, j)
    c = b + 1          # (i + 1, j + 1)
    d = a + 1          # (i, j + 1)
    f = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], axis=1).reshape(-1, 3)
    return v, f.astype(np.int64)


HEIGHTFIELD_TX_OFFSET = (97.31, 131.87, 25.0)   # x, y, height above the highest terrain point


def heightfield_scene(seed: int = 4, size: int = 128, pitch: float = 2.0, relief: float = 20.0,
                      boxes: int = 40, shape: tuple[int, int] = (256, 256), spacing: float | None = None) -> Scene:
    rng = np.random.default_rng(seed + 1000)
    heights = fbm_heights(seed, size, relief=relief)
    extent = (size - 1) * pitch
    soil = PhysicalMaterial(thickness_m=5.0, name="soil", **SOIL)
    walls = [PhysicalMaterial(thickness_m=0.2, name="concrete", **CONCRETE),
             PhysicalMaterial(thickness_m=0.02, name="glass", **GLASS),
             PhysicalMaterial(thickness_m=0.25, name="brick", **BRICK)]
    tv, tf = heightfield_mesh(heights, pitch)
    structures = [Structure(Mesh(torch.from_numpy(tv), torch.from_numpy(tf)), soil, name="terrain", surface_id=0)]
    tx_xy = np.array(HEIGHTFIELD_TX_OFFSET[:2])
    placed: list[tuple[float, float, float, float]] = []
    while len(placed) < boxes:
        w, l = rng.uniform(8.0, 20.0, size=2)
        x0 = rng.uniform(4.0, extent - 4.0 - w)
        y0 = rng.uniform(4.0, extent - 4.0 - l)
        box = (x0 - 3.0, y0 - 3.0, x0 + w + 3.0, y0 + l + 3.0)
        if any(box[0] < p[2] and p[0] < box[2] and box[1] < p[3] and p[1] < box[3] for p in placed):
            continue
        if box[0] <= tx_xy[0] <= box[2] and box[1] <= tx_xy[1] <= box[3]:
            continue
        placed.append(box)
        i0, i1 = int(x0 // pitch), int(math.ceil((x0 + w) / pitch))
        j0, j1 = int(y0 // pitch), int(math.ceil((y0 + l) / pitch))
        under = heights[i0:i1 + 1, j0:j1 + 1]
        base = float(under.min()) - 1.0
        top = float(under.max()) + float(rng.uniform(8.0, 30.0))
        v, f, _ = box_mesh(x0, y0, x0 + w, y0 + l, top - base)
        v[:, 2] += base
        mat = walls[int(rng.integers(0, len(walls)))]
        structures.append(Structure(Mesh(torch.from_numpy(v), torch.from_numpy(f)), mat, name=f"box-{len(placed)}",
                                    surface_id=len(placed)))
    top = float(heights.max())
    rows, cols = shape
    if spacing is None:
        spacing = extent / max(rows, cols)
    span_r, span_c = (rows - 1) * spacing, (cols - 1) * spacing
    origin = ((extent - span_r) / 2.0 + 0.137, (extent - span_c) / 2.0 + 0.219, top + 1.5)
    tx = (HEIGHTFIELD_TX_OFFSET[0], HEIGHTFIELD_TX_OFFSET[1], top + HEIGHTFIELD_TX_OFFSET[2])
    return Scene(structures=structures, endpoints=[transmitter(tx), grid(origin, shape, spacing)],
                 metadata={"imported_edge_policy": SimpleNamespace(edge_selection_mode="vertical_only")})


# eight materials of the scaled city: terrain soil, then seven wall materials
# (relative permittivity, conductivity S/m, slab thickness m)
CITY_MATERIALS = (
    ("soil", 15.0, 0.035, 5.0),
    ("concrete", 5.31, 0.0326, 0.2),
    ("glass", 6.27, 0.0043, 0.02),
    ("brick", 3.75, 0.038, 0.25),
    ("wood", 1.99, 0.0047, 0.05),
    ("plasterboard", 2.94, 0.0116, 0.015),
    ("marble", 7.074, 0.0055, 0.05),
    ("metal", 1.0, 1.0e7, 0.005),
)


def scaled_city(seed: int = 5, size: int = 256, pitch: float = 2.0, relief: float = 20.0, boxes: int = 400,
                tx_count: int = 128, side_range: tuple[int, int] = (64, 512)) -> SimpleNamespace:
    """Seeded city of ``boxes`` random boxes on a ``size`` x ``size`` fBm
    heightfield (``pitch`` m, ``relief`` m; 256 x 256 gives 130,050 terrain
    triangles and 4,800 box triangles), eight materials (``CITY_MATERIALS``),
    ``tx_count`` transmitters 3-30 m above the terrain outside every box, and
    one square receiver grid per transmitter whose side is drawn uniformly
    from ``side_range`` and which spans the terrain 1.5 m above its highest
    point. Diffraction edges follow the vertical-only policy.

    Returns ``scene`` (all structures and transmitters, plus the first
    transmitter's grid so the scene compiles) and the batched endpoint
    tables: ``tx`` [T, 3], ``sides`` [T], ``rx_count`` [T] (= sides^2) and
    ``rx`` [sum, 3] (each transmitter's grid, row-major).
    """
    rng = np.random.default_rng(seed + 2000)
    heights = fbm_heights(seed, size, relief=relief)
    extent = (size - 1) * pitch
    mats = [PhysicalMaterial(eps_r=e, sigma_e=s, thickness_m=t, name=n) for n, e, s, t in CITY_MATERIALS]
    tv, tf = heightfield_mesh(heights, pitch)
    structures = [Structure(Mesh(torch.from_numpy(tv), torch.from_numpy(tf)), mats[0], name="terrain",
                            surface_id=0)]
    placed: list[tuple[float, float, float, float]] = []
    attempts = 0
    while len(placed) < boxes:
        attempts += 1
        if attempts > 200_000:
            raise RuntimeError("scaled_city: could not place every box")
        w, l = rng.uniform(6.0, 16.0, size=2)
        x0 = rng.uniform(2.0, extent - 2.0 - w)
        y0 = rng.uniform(2.0, extent - 2.0 - l)
        box = (x0 - 1.5, y0 - 1.5, x0 + w + 1.5, y0 + l + 1.5)
        if any(box[0] < p[2] and p[0] < box[2] and box[1] < p[3] and p[1] < box[3] for p in placed):
            continue
        placed.append(box)
        i0, i1 = int(x0 // pitch), int(math.ceil((x0 + w) / pitch))
        j0, j1 = int(y0 // pitch), int(math.ceil((y0 + l) / pitch))
        under = heights[i0:i1 + 1, j0:j1 + 1]
        base = float(under.min()) - 1.0
        top = float(under.max()) + float(rng.uniform(6.0, 30.0))
        v, f, _ = box_mesh(x0, y0, x0 + w, y0 + l, top - base)
        v[:, 2] += base
        mat = mats[1 + int(rng.integers(0, len(mats) - 1))]
        structures.append(Structure(Mesh(torch.from_numpy(v), torch.from_numpy(f)), mat, name=f"box-{len(placed)}",
                                    surface_id=len(placed)))
    tx = []
    while len(tx) < tx_count:
        x, y = rng.uniform(4.0, extent - 4.0, size=2)
        if any(p[0] <= x <= p[2] and p[1] <= y <= p[3] for p in placed):
            continue
        i, j = int(round(x / pitch)), int(round(y / pitch))
        tx.append((x, y, float(heights[min(i, size - 1), min(j, size - 1)]) + float(rng.uniform(3.0, 30.0))))
    tx = np.asarray(tx, dtype=np.float32)
    sides = rng.integers(side_range[0], side_range[1] + 1, size=tx_count).astype(np.int64)
    z = float(heights.max()) + 1.5
    grids = []
    for s in sides:
        spacing = extent / float(s)
        axis = (np.arange(s, dtype=np.float32) + 0.5) * np.float32(spacing)
        gx, gy = np.meshgrid(axis, axis, indexing="ij")
        grids.append(np.stack([gx.reshape(-1), gy.reshape(-1), np.full(gx.size, z, np.float32)], axis=1))
    rx = np.concatenate(grids, axis=0).astype(np.float32)
    first = grid((0.5 * extent / sides[0], 0.5 * extent / sides[0], z), (int(sides[0]), int(sides[0])),
                 extent / float(sides[0]))
    scene = Scene(structures=structures, endpoints=[transmitter(tuple(map(float, p))) for p in tx] + [first],
                  metadata={"imported_edge_policy": SimpleNamespace(edge_selection_mode="vertical_only")})
    return SimpleNamespace(scene=scene, tx=tx, sides=sides, rx_count=sides * sides, rx=rx, extent=extent,
                           receiver_z=z, frequencies=(3.5e9, 28.0e9))


# thin walls of the batched path workload: (x0, y0, x1, y1) footprint in the
# street network, height, and the layer stack (thickness m, eps_r, sigma S/m)
STREET_WALLS = (
    ((52.0, 20.0, 52.0, 40.0), 8.0, ((0.12, 5.31, 0.0326),)),
    ((105.0, 97.0, 125.0, 97.0), 10.0, ((0.01, 6.27, 0.0043), (0.05, 1.0, 0.0), (0.01, 6.27, 0.0043))),
    ((187.0, 150.0, 187.0, 175.0), 6.0, ((0.02, 2.94, 0.0116), (0.1, 1.99, 0.0047))),
)


def street_wall(x0: float, y0: float, x1: float, y1: float, height: float, layers, surface_id: int) -> Structure:
    """Vertical thin-sheet wall on the segment (x0, y0)-(x1, y1), 0..height m,
    with a layered material (1-3 layers)."""
    v = np.array([[x0, y0, 0.0], [x1, y1, 0.0], [x0, y0, height], [x1, y1, height]], dtype=np.float32)
    f = np.array([[0, 1, 2], [1, 3, 2]], dtype=np.int64)
    mat = PhysicalMaterial(layers=tuple(MaterialLayer(thickness_m=t, eps_r=e, sigma_e=s) for t, e, s in layers),
                           name=f"wall-{surface_id}")
    return Structure(Mesh(torch.from_numpy(v), torch.from_numpy(f)), mat, name=f"street-wall-{surface_id}",
                     surface_id=surface_id)


def batched_path_workload(seed: int = 3, tx_count: int = 64) -> SimpleNamespace:
    """The urban canyon (seed 2) plus the three layered street walls of
    ``STREET_WALLS``; ``tx_count`` transmitters over the streets at 3-30 m,
    each with a street-point receiver set of log-uniform size 16..4096
    (``batched_transmitters``). Carrier 28 GHz; the channel frequency response
    is evaluated on 1,024 subcarriers spanning 100 MHz around it."""
    scene = urban_canyon_scene(seed=2, shape=(2, 2), spacing=2.0)
    walls = [street_wall(*seg, h, layers, 900 + k) for k, (seg, h, layers) in enumerate(STREET_WALLS)]
    scene = Scene(structures=list(scene.structures) + walls, endpoints=list(scene.endpoints))
    tx, sizes, rx = batched_transmitters(seed=seed, tx_count=tx_count)
    spacing = 100.0e6 / 1024
    subcarriers = 28.0e9 + (np.arange(1024) - 512) * spacing
    return SimpleNamespace(scene=scene, tx=tx, rx_count=sizes, rx=rx, frequency=28.0e9, subcarriers=subcarriers)
