"""CPU checks that pin the oracle's coupled reflection-diffraction paths
(oracle.cpp oracle_coupled_paths: R->D component 3, D->R component 4, D->D
component 7; reference interactions/coupled.py, native/channel/kernels/
coupled.cu) before the GPU tests compare against it.

* Image theory: over a perfectly conducting plane the reflected field of a
  dipole is the field of its image (position mirrored, tangential moment
  reversed). So the R->D field of a transmitter equals the first-order
  diffraction field of the image transmitter with the image polarization on
  the same scene without the plane, and the D->R field projected on the
  receiver polarization equals the diffraction field at the image receiver
  projected on the image polarization. The two sides share no code path: the
  coupled side runs the Jones reflection at the hit point and a UTD wedge fed
  with that field, the other side the plain first-order diffraction.
* Fermat: every coupled path's interaction points make the path length
  stationary (the derivatives along the reflecting plane and along each edge
  vanish), lie on their plane and edge lines, and the length is the sum of the
  legs. This pins the closed-form image / Keller constructions and the D->D
  bisection independently of how they are computed.

The D->D field amplitude (the spreading factor used for the second edge) has
no such closed-form check: parity unpinned (the reference's coupled field
kernels live in RayD, which is not in the reference tree).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from witwin.channel import workloads
from witwin.channel.scene.compiler import compile as compile_scene

F = 3.0e9
N = np.array([0.0, 0.0, 1.0])
NO_CHAINS = dict(max_depth=1, chains=np.zeros((0, 5), np.int32), chain_depth=np.zeros(0, np.int32))
CASES = [((0.0, -2.0, 1.0), (0.0, 2.0, 5.0)), ((5.5, 0.3, 3.0), (-1.0, 1.5, 2.5)), ((6.0, -0.5, 1.2), (0.0, 0.7, 3.5))]


def _edge_key(edges, i):
    return tuple(np.round([edges[k][i] for k in ("ax", "ay", "az", "dx", "dy", "dz")], 5))


def _scenes(tx, rx):
    a = compile_scene(workloads.coupled_wall_wedge_scene(tx=tx, rx=rx), reference_frequency_hz=F)
    b = compile_scene(workloads.coupled_wall_wedge_scene(tx=tx, rx=rx, ground=False), reference_frequency_hz=F)
    return a, b


@pytest.mark.parametrize("pol", [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.3, 0.8, 0.5)])
def test_reflect_then_diffract_is_the_image_source_diffraction(pol):
    checked = 0
    for tx, rx in CASES:
        checked += _rd_image_case(tx, rx, pol)
    assert checked >= 3


def _rd_image_case(tx, rx, pol) -> int:
    a, b = _scenes(tx, rx)
    oa, ob = oracle.Oracle(a), oracle.Oracle(b)
    by_key = {_edge_key(b.edges, i): i for i in range(b.edge_count)}
    p = np.asarray(pol) / np.linalg.norm(pol)
    T, X = np.asarray(tx), np.asarray(rx)
    image, p_image = T - 2 * np.dot(T, N) * N, 2 * np.dot(p, N) * N - p
    rd = oa.coupled_paths(tx=T, rx=X, tx_pol=p, rx_pol=p, frequency=F, kinds=1)
    direct = ob.pair_paths(tx=image, rx=X, tx_pol=p_image, rx_pol=p, frequency=F, components=4, **NO_CHAINS)
    checked = 0
    for path in rd:
        if abs(path.pos[0][2]) > 1e-9:
            continue  # reflected by a wedge face, not by the plane z = 0
        e = by_key.get(_edge_key(a.edges, path.edge))
        match = [q for q in direct if q.edge == e]
        if e is None or not match:
            continue
        c_rd, c_img = complex(*path.coefficient), complex(*match[0].coefficient)
        # the effective conductivity of the "perfect" conductor leaves ~5e-5
        assert abs(c_rd - c_img) <= 2e-4 * abs(c_img)
        assert path.length == pytest.approx(match[0].length, rel=1e-12)
        checked += 1
    return checked


def test_diffract_then_reflect_is_the_image_receiver_diffraction():
    checked = 0
    for tx, rx in CASES:
        checked += _dr_image_case(tx, rx)
    assert checked >= 3


def _dr_image_case(tx, rx) -> int:
    a, b = _scenes(tx, rx)
    oa, ob = oracle.Oracle(a), oracle.Oracle(b)
    by_key = {_edge_key(b.edges, i): i for i in range(b.edge_count)}
    p = np.array([0.3, 0.8, 0.5]) / np.linalg.norm([0.3, 0.8, 0.5])
    T, X = np.asarray(tx), np.asarray(rx)
    x_image, p_image = X - 2 * np.dot(X, N) * N, 2 * np.dot(p, N) * N - p
    dr = oa.coupled_paths(tx=T, rx=X, tx_pol=p, rx_pol=p, frequency=F, kinds=2)
    direct = ob.pair_paths(tx=T, rx=x_image, tx_pol=p, rx_pol=p_image, frequency=F, components=4, **NO_CHAINS)
    checked = 0
    for path in dr:
        if abs(path.pos[1][2]) > 1e-9:
            continue
        e = by_key.get(_edge_key(a.edges, path.edge))
        match = [q for q in direct if q.edge == e]
        if e is None or not match:
            continue
        c_dr, c_img = complex(*path.coefficient), complex(*match[0].coefficient)
        assert abs(c_dr - c_img) <= 2e-4 * abs(c_img)
        checked += 1
    return checked


def _unit(v):
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("tx,rx", CASES)
def test_coupled_paths_are_stationary(tx, rx):
    """Fermat: d(length) = 0 along the plane (R->D, D->R) and along every edge."""
    a, _ = _scenes(tx, rx)
    o = oracle.Oracle(a)
    T, X = np.asarray(tx), np.asarray(rx)
    e = a.edges
    E = a.edge_count
    A = np.stack([e["ax"], e["ay"], e["az"]], 1).astype(np.float64)
    D = np.stack([e["dx"], e["dy"], e["dz"]], 1).astype(np.float64)
    kinds_seen = set()
    for path in o.coupled_paths(tx=T, rx=X, frequency=F):
        p0, p1 = np.array(path.pos[0]), np.array(path.pos[1])
        kinds_seen.add(path.component)
        if path.component == 7:
            e1, e2 = divmod(path.edge, E)
            legs = [T, p0, p1, X]
            for q, k in ((p0, e1), (p1, e2)):
                off = q - A[k]
                assert np.linalg.norm(off - D[k] * np.dot(off, D[k])) < 1e-9  # on the edge line
            g1 = np.dot(D[e1], _unit(p0 - T) + _unit(p0 - p1))
            g2 = np.dot(D[e2], _unit(p1 - p0) + _unit(p1 - X))
            assert abs(g1) < 1e-6 and abs(g2) < 1e-6
        else:
            R, Q = (p0, p1) if path.component == 3 else (p1, p0)
            k = path.edge
            plane = a.group_plane[a.face_group[path.prim[0]]]
            n = np.asarray(plane[:3], np.float64)
            assert abs(np.dot(n, R) - plane[3]) < 1e-9  # on the reflecting plane
            off = Q - A[k]
            assert np.linalg.norm(off - D[k] * np.dot(off, D[k])) < 1e-9
            legs = [T, R, Q, X] if path.component == 3 else [T, Q, R, X]
            # stationarity at R within the plane and at Q along the edge
            prev_r, next_r = (T, Q) if path.component == 3 else (Q, X)
            prev_q, next_q = (R, X) if path.component == 3 else (T, R)
            grad_r = _unit(R - prev_r) + _unit(R - next_r)
            assert np.linalg.norm(grad_r - n * np.dot(grad_r, n)) < 1e-6
            assert abs(np.dot(D[k], _unit(Q - prev_q) + _unit(Q - next_q))) < 1e-6
        total = sum(np.linalg.norm(legs[i + 1] - legs[i]) for i in range(3))
        assert path.length == pytest.approx(total, rel=1e-12)
    assert kinds_seen == {3, 4, 7}


@pytest.mark.parametrize("ad_mode", ["vjp", "jvp"])
def test_coupled_with_ad_or_export_raises_before_any_launch(ad_mode):
    from witwin.channel.deterministic import Config, solve

    scene = workloads.coupled_wall_wedge_scene()
    comps = {"los", "reflection", "diffraction"}
    with pytest.raises(NotImplementedError, match="forward only"):
        solve(scene, Config(components=comps, max_depth=2, coupled_paths=True, ad_mode=ad_mode),
              reference_frequency_hz=F)
    with pytest.raises(NotImplementedError, match="radiomap only"):
        solve(scene, Config(components=comps, max_depth=2, coupled_paths=True, export_paths=True),
              reference_frequency_hz=F)
    with pytest.raises(RuntimeError, match="max_depth >= 2"):
        Config(components=comps, max_depth=1, coupled_paths=True)


def test_device_code_on_the_cpu_matches_the_oracle(tmp_path):
    """The coupled candidate code of the kernels (wc_coupled.h, built as host
    C++ through the shim in cpu_geo/hip) over a 14 x 14 receiver grid around
    the wall-wedge: the same path set as the oracle on every cell whose oracle
    path set is stable under 0.2 mm receiver moves, and the same summed field."""
    import shutil
    import subprocess
    from pathlib import Path

    from test_cpu_oracle_and_host import _sah_tables

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    here = Path(__file__).resolve().parent
    exe = tmp_path / "coupled_check"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{here / 'cpu_geo'}", str(here / "cpu_geo" / "coupled_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    scene = workloads.coupled_wall_wedge_scene(tx=(5.5, 0.3, 3.0), grid_shape=(14, 14), grid_origin=(-3.0, -3.2, 2.6),
                                               grid_spacing=0.6)
    c = compile_scene(scene, reference_frequency_hz=F)
    d = tmp_path / "scene"
    d.mkdir()
    nodes, sorted_tris, face_tris = _sah_tables(c)
    nodes.tofile(d / "nodes.bin")
    sorted_tris.astype(np.float32).tofile(d / "tris.bin")
    face_tris.astype(np.float32).tofile(d / "faces.bin")
    c.face_group.astype(np.int32).tofile(d / "fgroup.bin")
    c.face_material.astype(np.int32).tofile(d / "fmat.bin")
    c.group_plane.astype(np.float64).tofile(d / "plane.bin")
    c.group_face_offset.astype(np.int32).tofile(d / "goff.bin")
    c.group_faces.astype(np.int32).tofile(d / "gfaces.bin")
    m = c.materials
    np.stack([m.eps_r, m.sigma_e, m.mu_r, m.gain, m.thickness_m], 1).astype(np.float32).tofile(d / "mats.bin")
    c.edges.tofile(d / "edges.bin")
    tx = c.endpoints.tx_position[0]
    np.array(list(tx) + [0, 0, 1, 0, 0, 1, 1.0, F], np.float64).tofile(d / "ends.bin")
    rx = np.asarray(c.endpoints.rx_position, np.float32)
    rx.tofile(d / "rx.bin")
    lines = subprocess.run([str(exe), str(d)], capture_output=True, text=True, check=True, timeout=300).stdout
    cells = {}
    for line in lines.splitlines():
        t = line.split()
        if t[0] == "cell":
            cells[int(t[1])] = (int(t[2]), complex(float(t[4]), float(t[5])))
    o = oracle.Oracle(c)
    moves = np.concatenate([np.eye(3), -np.eye(3)]) * 2e-4
    ambiguous = 0
    for cell in range(len(rx)):
        x = rx[cell].astype(np.float64)
        paths = o.coupled_paths(tx=tx, rx=x, frequency=F)
        sig = sorted((p.component, p.edge, p.prim[0]) for p in paths)
        n, field = cells[cell]
        if n != len(paths):
            moved = [sorted((p.component, p.edge, p.prim[0]) for p in o.coupled_paths(tx=tx, rx=x + dm, frequency=F))
                     for dm in moves]
            assert any(mv != sig for mv in moved), f"cell {cell}: stable oracle set, {len(paths)} vs {n} paths"
            ambiguous += 1
            continue
        ref = sum(complex(*p.path_field) for p in paths)
        assert abs(field - ref) <= 1e-4 * abs(ref) + 1e-12
    assert ambiguous <= 6
