"""The conservative culls checked on the CPU against the exact float tests,
on the device sources themselves (compiled as host C++ through a shim).

Reflection chains per tile: a chain that chain_reaches_box drops for an 8 x 8
tile of receivers must not back-project into faces of all its groups for any
receiver of that tile.

The 4-wide tree walk: any-hit and closest-hit walks over the host-built tree
must give the answers of a test against every triangle, for random segments
and for segments aimed through shared vertices and edges (where the widened
box test and the face-id tie-break matter); closest hits may differ only
where two surfaces meet within the triangle test's rounding of the distance.

Transmitter-leg shadow bins: every bin a triangle sets must be blocked by
that triangle for legs sampled across the bin (float watertight test, the
leg test's end bias); the float pre-test that skips triangles whose in-plane
hull misses the (transmitter, edge) fan triangle must never skip a triangle
that sets a bin."""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from witwin.channel import workloads
from witwin.channel.scene.compiler import compile as compile_scene

HERE = Path(__file__).resolve().parent


def _dump(scene, txs: np.ndarray, out: Path, frequency: float) -> None:
    compiled = compile_scene(scene, reference_frequency_hz=frequency)
    tri = np.zeros((compiled.face_count, 12), np.float32)
    for k in range(3):
        tri[:, 4 * k:4 * k + 3] = compiled.vertices[compiled.faces[:, k]]
    tri[:, 3] = compiled.face_group.astype(np.int32).view(np.float32)
    out.mkdir(parents=True, exist_ok=True)
    tri.tofile(out / "tris.bin")
    compiled.edges.view(np.uint8).tofile(out / "edges.bin")
    np.asarray(txs, np.float32).reshape(-1, 3).tofile(out / "txs.bin")


def _build(tmp_path_factory, name: str) -> Path:
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("cull") / name
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{HERE / 'cpu_geo'}", str(HERE / "cpu_geo" / f"{name}.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    return exe


@pytest.fixture(scope="module")
def checker(tmp_path_factory) -> Path:
    return _build(tmp_path_factory, "shadow_check")


@pytest.fixture(scope="module")
def chain_checker(tmp_path_factory) -> Path:
    return _build(tmp_path_factory, "chain_cull_check")


def _dump_chains(scene, out: Path, seed: int, depth1: int, depth2: int, discovered_depth: int = 0) -> None:
    """Scene tables, a random sample of depth-1 and depth-2 group chains plus
    (``discovered_depth`` > 0) the distinct prefixes up to that depth of the
    FP64 oracle's discovery rays, with their transmitter images, the group
    hulls and the scene's square receiver grid."""
    from witwin.channel.scene.compiler import group_hulls

    compiled = compile_scene(scene, reference_frequency_hz=3.5e9)
    out.mkdir(parents=True, exist_ok=True)
    g = compiled.group_count
    rng = np.random.default_rng(seed)
    one = rng.choice(g, size=min(g, depth1), replace=False)
    two = rng.integers(0, g, size=(depth2, 2))
    two = two[two[:, 0] != two[:, 1]]
    chains = np.full((len(one) + len(two), 5), -1, np.int32)
    chains[:len(one), 0] = one
    chains[len(one):, :2] = two
    tx = compiled.endpoints.tx_position[0].astype(np.float64)
    if discovered_depth:
        import oracle

        rays = oracle.Oracle(compiled).discover_all(tx, 16384, discovered_depth)
        found = []
        for d in range(1, discovered_depth + 1):
            pref = rays[:, :d]
            pref = np.unique(pref[(pref >= 0).all(axis=1)], axis=0)
            block = np.full((len(pref), 5), -1, np.int32)
            block[:, :d] = pref
            found.append(block)
        chains = np.concatenate([chains] + found, axis=0)
    hull_offset, hull_points = group_hulls(compiled.vertices, compiled.faces, compiled.group_face_offset,
                                           compiled.group_faces, compiled.group_plane)
    hull_offset.astype(np.int32).tofile(out / "hulloff.bin")
    hull_points.astype(np.float64).tofile(out / "hullpts.bin")
    n, off = compiled.group_plane[:, :3], compiled.group_plane[:, 3]
    images = np.zeros((len(chains), 5, 3))
    p = np.repeat(tx[None], len(chains), 0)
    for j in range(5):
        grp = chains[:, j]
        ok = grp >= 0
        gg = np.where(ok, grp, 0)
        d = (p * n[gg]).sum(1) - off[gg]
        p = np.where(ok[:, None], p - 2.0 * d[:, None] * n[gg], p)
        images[:, j] = np.where(ok[:, None], p, 0.0)
    tri = np.zeros((compiled.face_count, 12), np.float32)
    for k in range(3):
        tri[:, 4 * k:4 * k + 3] = compiled.vertices[compiled.faces[:, k]]
    tri[:, 3] = compiled.face_group.astype(np.int32).view(np.float32)
    rx = np.asarray(compiled.endpoints.rx_position, np.float32)
    side = int(round(np.sqrt(len(rx))))
    compiled.group_plane.astype(np.float64).tofile(out / "plane.bin")
    compiled.group_face_offset.astype(np.int32).tofile(out / "goff.bin")
    compiled.group_faces.astype(np.int32).tofile(out / "gfaces.bin")
    tri.tofile(out / "tris.bin")
    chains.tofile(out / "chains.bin")
    images.tofile(out / "images.bin")
    rx.tofile(out / "rx.bin")
    tx.astype(np.float32).tofile(out / "tx.bin")
    np.array([g, compiled.face_count, len(chains), side], np.int32).tofile(out / "counts.bin")


def _run(exe: Path, data: Path) -> dict[str, int]:
    proc = subprocess.run([str(exe), str(data)], capture_output=True, text=True, timeout=600)
    counts = {k: int(v) for k, v in re.findall(r"(\w+) (\d+)", proc.stdout)}
    assert proc.returncode == 0, proc.stdout + proc.stderr
    return counts


def test_shadow_bins_and_pretest_on_urban_canyon(checker, tmp_path):
    scene = workloads.urban_canyon_scene(seed=2, shape=(2, 2))
    tx_batch, _, _ = workloads.batched_transmitters(seed=3, tx_count=2)
    txs = np.concatenate([[[53.231, 52.013, 25.0]], tx_batch])
    _dump(scene, txs, tmp_path / "canyon", 28e9)
    c = _run(checker, tmp_path / "canyon")
    assert c["violations"] == 0 and c["skipped_with_bins"] == 0
    assert c["set_bins"] > 1000 and c["skipped"] > c["tests"] // 2


def test_shadow_bins_and_pretest_on_heightfield(checker, tmp_path):
    scene = workloads.heightfield_scene(seed=4, size=64, boxes=16, shape=(2, 2))
    _dump(scene, scene.endpoints[0].position.numpy().reshape(1, 3), tmp_path / "terrain", 3.5e9)
    c = _run(checker, tmp_path / "terrain")
    assert c["violations"] == 0 and c["skipped_with_bins"] == 0
    assert c["set_bins"] > 0 and c["skipped"] > c["tests"] // 2


@pytest.mark.parametrize("which", ["canyon", "heightfield"])
def test_chain_tile_cull_drops_no_valid_chain(chain_checker, tmp_path, which):
    if which == "canyon":
        scene = workloads.urban_canyon_scene(seed=2, shape=(24, 24), spacing=14.0)
    else:
        scene = workloads.heightfield_scene(seed=4, size=40, boxes=10, shape=(24, 24))
    _dump_chains(scene, tmp_path / which, seed=0, depth1=1500, depth2=1500, discovered_depth=4)
    c = _run(chain_checker, tmp_path / which)
    assert c["wrong"] == 0
    assert c["culled"] > c["kept"] // 4 and c["kept_valid"] > 0
    # the receiver cones bound most chains and drop (tile, chain) pairs the
    # last-bounce face test keeps
    assert c["bounded"] > c["chains"] // 2 and c["cone_only"] > 0


@pytest.fixture(scope="module")
def walk_checker(tmp_path_factory) -> Path:
    return _build(tmp_path_factory, "walk_check")


def test_wide_walk_matches_every_triangle(walk_checker, tmp_path):
    from test_cpu_oracle_and_host import _sah_tables

    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    nodes, sorted_tris, _ = _sah_tables(compiled)
    rng = np.random.default_rng(21)
    lo, hi = compiled.vertices.min(axis=0), compiled.vertices.max(axis=0)
    n = 3000
    a = rng.uniform(lo - 5.0, hi + 5.0, (n, 3))
    b = rng.uniform(lo - 5.0, hi + 5.0, (n, 3))
    # segments through vertices and edge midpoints, in random directions
    tri = compiled.vertices[compiled.faces].astype(np.float64)
    through = np.concatenate([tri.reshape(-1, 3), 0.5 * (tri[:, 0] + tri[:, 1]), 0.5 * (tri[:, 1] + tri[:, 2])])
    pick = through[rng.integers(0, len(through), 3000)]
    dirs = rng.normal(size=(3000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    reach = rng.uniform(2.0, 60.0, (3000, 1))
    segs = np.concatenate([np.concatenate([a, b], 1), np.concatenate([pick - dirs * reach, pick + dirs * reach], 1)])
    out = tmp_path / "walk"
    out.mkdir()
    nodes.tofile(out / "nodes.bin")
    sorted_tris.astype(np.float32).tofile(out / "tris.bin")
    segs.astype(np.float32).tofile(out / "segs.bin")
    c = _run(walk_checker, out)
    assert c["any_wrong"] == 0 and c["closest_wrong"] == 0
    assert c["hit"] > 1000 and c["near_ties"] <= c["segments"] // 1000
