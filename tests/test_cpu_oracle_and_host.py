"""CPU tests: the FP64 oracle against golden vectors and closed forms, the host
compile logic, and the C ABI of libwcmi.so (load + exported symbols only)."""

from __future__ import annotations

import cmath
import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from witwin.channel.propagation.engine import (cell_order_table, exhaustive_sequences, is_exhaustive,
                                               locality_permutation, make_tiles)
from witwin.channel.scene.compiler import build_edges, compile as compile_scene, coplanar_face_groups
from witwin.channel import workloads
from witwin.core import Mesh, PhysicalMaterial, Scene, Structure

ROOT = Path(__file__).resolve().parents[1]
EPS0 = 8.8541878128e-12
C0 = 299792458.0


# ----------------------------------------------------------------- oracle pins

def test_slab_known_answer():
    """Reference tests/kernels/test_field_transport.py:116-159 (normal incidence,
    3.5 GHz, eps 4, sigma 0.025, gain 0.8, t 0.12 m, L 3 m)."""
    f = 3.5e9
    omega = 2 * math.pi * f
    lam = C0 / f
    eta = complex(4.0, -0.025 / (omega * EPS0))
    root = cmath.sqrt(eta)
    r0 = (1 - root) / (1 + root)
    ph = cmath.exp(-2j * 2 * math.pi * 0.12 / lam * root)
    slab = 0.8 * r0 * (1 - ph) / (1 - r0 * r0 * ph)
    expected = lam / (4 * math.pi * 3.0) * cmath.exp(-2j * math.pi * 3.0 / lam) * slab
    row = oracle.reflection_sequence([0, 0, 1.0], [0, 0, 2.0], [[0, 0, 0.0]], [[0, 0, 1.0]], 1.0, [1, 0, 0],
                                     [1, 0, 0], [4.0], [0.025], [1.0], [0.8], [0.12], f)
    assert abs(row["coefficient"] - expected) <= 1e-8 * abs(expected)


def test_slab_limits():
    # thick lossy slab -> half-space Fresnel; PEC -> -1 / +1; zero thickness -> 0
    te, tm = oracle.slab_fresnel(0.6, 5.31, 0.0326, 1.0, 1.0, 5.0, 3.5e9)
    eta = complex(5.31, -0.0326 / (2 * math.pi * 3.5e9 * EPS0))
    root = cmath.sqrt(eta - (1 - 0.36))
    assert abs(te - (0.6 - root) / (0.6 + root)) < 1e-10
    assert abs(tm - (eta * 0.6 - root) / (eta * 0.6 + root)) < 1e-10
    te, tm = oracle.slab_fresnel(0.3, 1.0, 1e9, 1.0, 1.0, 0.1, 3e9)
    assert abs(te + 1) < 1e-4 and abs(tm - 1) < 1e-4
    te, tm = oracle.slab_fresnel(0.3, 4.0, 0.01, 1.0, 1.0, 0.0, 3e9)
    assert abs(te) < 1e-15 and abs(tm) < 1e-15
    # Brewster null of the TM interface coefficient (lossless, thick limit via large loss-free phase average)
    eps = 4.0
    ct_b = math.cos(math.atan(math.sqrt(eps)))
    eta = complex(eps, 0.0)
    root = cmath.sqrt(eta - (1 - ct_b * ct_b))
    assert abs((eta * ct_b - root) / (eta * ct_b + root)) < 1e-12


def test_free_space_closed_form():
    f = 3.0e9
    row = oracle.free_space([0, 0, 0], [10.0, 0, 0], 4.0, [0, 0, 1], [0, 0, 1], f)
    lam = C0 / f
    expected = lam / (4 * math.pi * 10.0) * cmath.exp(-2j * math.pi * 10.0 / lam)
    assert abs(row["coefficient"] - expected) < 1e-12 * abs(expected) + 1e-18
    assert abs(row["path_field"] - 2.0 * expected) < 1e-12
    # dipole pattern: p_rx.p_tx - (p_tx.k)(p_rx.k) (reference tests/support/reference_channel.py:63-77)
    row = oracle.free_space([0, 0, 0], [3.0, 4.0, 1.0], 1.0, [0, 0, 1], [0, 0, 1], f)
    k = np.array([3.0, 4.0, 1.0]) / math.sqrt(26.0)
    dip = (1 - k[2] * k[2]) ** 2
    assert abs(row["path_gain"] - (lam / (4 * math.pi * math.sqrt(26.0))) ** 2 * dip) < 1e-12 * row["path_gain"]


def test_fresnel_integrals_match_scipy():
    from scipy.special import fresnel as sp_fresnel

    for x in np.concatenate([np.linspace(-6, 6, 241), [0.01, 1.49, 1.51, 3.99, 4.01, 10, 40]]):
        s_ref, c_ref = sp_fresnel(x)
        c, s = oracle.fresnel(float(x))
        assert abs(c - c_ref) < 2e-9 and abs(s - s_ref) < 2e-9, x


def test_transition_function_limits():
    # F(X) -> 1 for large X; small-X form F ~ sqrt(pi X) e^{j pi/4}
    big = oracle.transition(200.0)
    assert abs(big - (1 + 0.5j / 200.0)) < 1e-4
    small = oracle.transition(1e-6)
    assert abs(small - math.sqrt(math.pi * 1e-6) * cmath.exp(1j * math.pi / 4)) < 1e-5


def _utd_scalar_field(n, phi, phi_p, k, rho, rho_p, r0=-1.0, rn=-1.0):
    """2-D (normal incidence) soft diffracted field of a unit cylindrical-like
    source with the UTD coefficient; used only for continuity checks."""
    L = rho * rho_p / (rho + rho_p)
    ds, _ = oracle.utd(n, phi, phi_p, 1.0, k, L, r0s=r0, rns=rn)
    spread = math.sqrt(rho_p / (rho * (rho + rho_p)))
    return ds * spread * cmath.exp(-1j * k * (rho + rho_p)) / rho_p


@pytest.mark.parametrize("n", [1.5, 2.0])
def test_utd_compensates_incident_shadow_boundary(n):
    """Across the incident shadow boundary phi = pi + phi' the diffracted field
    jumps by exactly the (removed) incident field: GO + diffraction is continuous.
    A sign error in any cot term breaks this."""
    k = 2 * math.pi / 0.1
    rho_p, rho = 7.0, 5.0
    phi_p = 0.6
    isb = math.pi + phi_p
    d = 1e-7
    lit = _utd_scalar_field(n, isb - d, phi_p, k, rho, rho_p)
    shadow = _utd_scalar_field(n, isb + d, phi_p, k, rho, rho_p)
    incident = cmath.exp(-1j * k * (rho + rho_p)) / (rho + rho_p) / rho_p * rho_p  # unit amplitude at distance
    # GO field at the boundary: incident wave at distance rho + rho_p (spherical spreading ~ 1/(rho+rho_p))
    go = cmath.exp(-1j * k * (rho + rho_p)) / (rho + rho_p)
    jump = shadow - lit
    assert abs(jump - go) < 2e-3 * abs(go), (jump, go)
    del incident


def test_utd_compensates_reflection_shadow_boundary():
    k = 2 * math.pi / 0.1
    rho_p, rho = 7.0, 5.0
    phi_p = 0.6
    n = 1.5
    rsb = math.pi - phi_p
    d = 1e-7
    before = _utd_scalar_field(n, rsb - d, phi_p, k, rho, rho_p, r0=-1.0)
    after = _utd_scalar_field(n, rsb + d, phi_p, k, rho, rho_p, r0=-1.0)
    go = -1.0 * cmath.exp(-1j * k * (rho + rho_p)) / (rho + rho_p)  # reflected (soft: R = -1)
    assert abs((after - before) - go) < 2e-3 * abs(go)


def test_stack_transmission_identities():
    f = 5e9
    # vacuum layer is transparent
    t = oracle.stack_t(0.7, [0.3], [1.0], [0.0], [1.0], f, 0)
    assert abs(abs(t) - 1.0) < 1e-12
    # lossless single layer: |r|^2 + |t|^2 = 1 (same medium both sides); compare with slab r
    for pol in (0, 1):
        t = oracle.stack_t(0.8, [0.125], [4.0], [0.0], [1.0], f, pol)
        te, tm = oracle.slab_fresnel(0.8, 4.0, 0.0, 1.0, 1.0, 0.125, f)
        r = te if pol == 0 else tm
        assert abs(abs(r) ** 2 + abs(t) ** 2 - 1.0) < 1e-10
    # two identical layers equal one layer of double thickness
    t2 = oracle.stack_t(0.5, [0.1, 0.1], [3.0, 3.0], [0.01, 0.01], [1.0, 1.0], f, 0)
    t1 = oracle.stack_t(0.5, [0.2], [3.0], [0.01], [1.0], f, 0)
    assert abs(t2 - t1) < 1e-12


def _two_wall_scene():
    def wall(v, f, mat, sid):
        return Structure(Mesh(torch.tensor(v), torch.tensor(f)), mat, surface_id=sid)

    wx = wall([[2.0, -1.0, 0.0], [2.0, 3.0, 0.0], [2.0, -1.0, 2.0], [2.0, 3.0, 2.0]], [[0, 1, 2], [1, 3, 2]],
              PhysicalMaterial(eps_r=3.0, sigma_e=0.005), 10)
    wy = wall([[0.0, 2.0, 0.0], [3.0, 2.0, 0.0], [0.0, 2.0, 2.0], [3.0, 2.0, 2.0]], [[0, 2, 1], [1, 2, 3]],
              PhysicalMaterial(eps_r=4.0, sigma_e=0.01), 11)
    return Scene(structures=[wx, wy], endpoints=[workloads.transmitter((0.0, 0.0, 1.0)),
                                                 workloads.receiver((0.0, 1.0, 1.0))])


def _exhaustive_chains(groups, depth):
    chains, depths = [], []
    for d in range(1, depth + 1):
        s = exhaustive_sequences(groups, d)
        pad = np.full((s.shape[0], 5), -1, np.int32)
        pad[:, :d] = s
        chains.append(pad)
        depths.append(np.full(s.shape[0], d, np.int32))
    return np.concatenate(chains), np.concatenate(depths)


def test_oracle_multibounce_golden_vectors():
    """Reference tests/deterministic/test_reflection_multibounce.py:179-215."""
    c = compile_scene(_two_wall_scene(), reference_frequency_hz=3e9)
    chains, depths = _exhaustive_chains(c.group_count, 2)
    paths = oracle.Oracle(c).pair_paths(tx=(0, 0, 1), rx=(0, 1, 1), frequency=3e9, components=2, max_depth=2,
                                        chains=chains, chain_depth=depths)
    assert [p.depth for p in paths] == [1, 1, 2]
    assert [list(p.prim)[:2] for p in paths] == [[0, -1], [2, -1], [1, 2]]
    # golden lengths are quoted to 10 digits; tolerance as in the reference test
    np.testing.assert_allclose([p.length for p in paths], [4.123105526, 3.0, 5.0], rtol=1e-5, atol=1e-6)
    assert abs(paths[0].length - math.sqrt(17.0)) < 1e-12


def test_oracle_transmission_vacuum_wall_equals_los():
    """Reference tests/deterministic/test_transmission.py:88-120: a vacuum wall
    (eps_r = 1, sigma = 0) leaves the line-of-sight field unchanged."""
    vac = PhysicalMaterial(eps_r=1.0, sigma_e=0.0, thickness_m=0.3)
    v = np.array([[2.0, -4, -4], [2.0, 4, -4], [2.0, -4, 4], [2.0, 4, 4]], np.float32)
    wall = Structure(Mesh(torch.from_numpy(v), torch.tensor([[0, 1, 2], [1, 3, 2]])), vac, surface_id=1)
    scene = Scene(structures=[wall], endpoints=[workloads.transmitter((0.0, 0.0, 0.0)),
                                                workloads.receiver((5.0, 1.0, 0.5))])
    c = compile_scene(scene, reference_frequency_hz=3e9)
    o = oracle.Oracle(c)
    paths = o.pair_paths(tx=(0, 0, 0), rx=(5, 1, 0.5), frequency=3e9, components=9, max_depth=1,
                         chains=np.zeros((0, 5), np.int32), chain_depth=np.zeros(0, np.int32))
    assert [p.component for p in paths] == [5]
    ref = oracle.free_space([0, 0, 0], [5, 1, 0.5], 1.0, [0, 0, 1], [0, 0, 1], 3e9)
    # the vacuum wall's stack phase exp(-j k d cos) cancels the carrier shortening
    assert abs(complex(*paths[0].coefficient) - ref["coefficient"]) < 1e-4 * abs(ref["coefficient"])


def test_oracle_canonical_sort_contract():
    """Reference sort-order vector (test_reflection_multibounce.py:25-80): rows
    with sequences [[4,2],[4,1],[3,9]] sort to [[3,9],[4,1],[4,2]]."""
    seqs = np.array([[4, 2], [4, 1], [3, 9]])
    order = sorted(range(3), key=lambda i: tuple(seqs[i]))
    assert seqs[order].tolist() == [[3, 9], [4, 1], [4, 2]]


def _compact_rows(spec):
    """Rows from (depth, component, sequence, primitive, edge, length) tuples."""
    return {"depth": np.array([r[0] for r in spec], np.int32),
            "component_id": np.array([r[1] for r in spec], np.int32),
            "primitive_sequence": np.array([r[2] for r in spec], np.int32),
            "primitive_id": np.array([r[3] for r in spec], np.int32),
            "edge_id": np.array([r[4] for r in spec], np.int32),
            "path_length_m": np.array([r[5] for r in spec], np.float32)}


def test_oracle_canonical_compact_known_answers():
    """Hand-worked canonical compact of three pairs (reference
    evaluated_paths.cu:126-320, 654-700): the sort-order vector [[4,2],[4,1],
    [3,9]] -> [[3,9],[4,1],[4,2]] after the line-of-sight row; a longer
    duplicate of [4,1] dropped for the shorter one; two equal-length
    duplicates keep the first in sorted order; diffraction rows are identified
    by their edge; a NaN length leaves its group without a winner; an empty
    pair in between."""
    spec = [
        # pair 0 (rows 0-4)
        (2, 1, [4, 2], 2, -1, 10.0),
        (2, 1, [4, 1], 1, -1, 11.0),
        (2, 1, [3, 9], 9, -1, 12.0),
        (0, 0, [-1, -1], -1, -1, 5.0),
        (2, 1, [4, 1], 1, -1, 10.5),
        # pair 2 (rows 5-9); pair 1 is empty
        (1, 2, [6, -1], 6, 7, 3.0),
        (1, 2, [6, -1], 6, 8, 4.0),
        (1, 1, [6, -1], 6, -1, 3.5),
        (1, 1, [6, -1], 6, -1, 3.5),
        (1, 5, [3, -1], 3, -1, float("nan")),
    ]
    rows = _compact_rows(spec)
    out = oracle.canonical_compact(rows, np.array([0, 5, 5, 10]))
    # pair 0 sorted: LoS (3), [3,9] (2), [4,1] (1), [4,1] (4), [4,2] (0); rows 1 and 4 share an identity
    assert out["perm"][:5].tolist() == [3, 2, 1, 4, 0]
    assert out["keep"][:5].tolist() == [1, 1, 0, 1, 1]
    # pair 2 sorted: reflection d1 (7, 8: equal lengths, first kept), diffraction edges 7 then 8, transmission (NaN)
    assert out["perm"][5:].tolist() == [7, 8, 5, 6, 9]
    assert out["keep"][5:].tolist() == [1, 0, 1, 1, 0]
    assert out["kept_offset"].tolist() == [0, 4, 4, 7]
    assert out["source_row"].tolist() == [3, 2, 4, 0, 7, 5, 6]
    assert out["dropped"] == 0
    capped = oracle.canonical_compact(rows, np.array([0, 5, 5, 10]), cap=2)
    assert capped["kept_offset"].tolist() == [0, 2, 2, 4]
    assert capped["source_row"].tolist() == [3, 2, 7, 5]
    assert capped["dropped"] == 3


# reference tests/ad/test_deterministic_accum_ad.py:38-82: 12 rows over a 2 x 2
# grid covering every slot (a same-slot collision, transmission, scattering,
# two coupled rows, an out-of-range tx and rx) and the frozen SHA-256 of the
# six outputs for each (dtype, coherent)
ACCUM_TX = (0, 0, 0, 1, 1, 0, -1, 0, 0, 0, 1, 1)
ACCUM_RX = (0, 0, 1, 1, 0, 1, 0, 3, 0, 0, 1, 1)
ACCUM_CID = (0, 1, 1, 2, 0, 5, 1, 2, 0, 6, 3, 4)
ACCUM_SHA256 = {
    ("float32", True): "29c4e0ef1406dfa4e145967dc088535730e59e9538fee1ab4c6fcfe42a380fbe",
    ("float32", False): "d68266d382aeb6ae7fdf6479660e29eea039065bb7645eeafcc7aad25ebf303d",
    ("float64", True): "74d28d783c3f5d6613a305eeb66f63644e3373000a6d36272b22409813bfe7e7",
    ("float64", False): "70ca8b66f1bbc6a97fb39582ca7cd02cc1037a55a9e9891e4b3a5261453b669e",
}
ACCUM_FIELDS = ("power_total", "field_total_real", "field_total_imag", "component_power", "component_field_real",
                "component_field_imag")


def accum_fixture():
    """(gain, real, imag) float64 rows of the reference fixture (CPU generator
    seeded 3: rand + 0.25, randn, randn)."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(3)
    n = len(ACCUM_TX)
    gain = torch.rand((n,), generator=g, dtype=torch.float64) + 0.25
    real = torch.randn((n,), generator=g, dtype=torch.float64)
    imag = torch.randn((n,), generator=g, dtype=torch.float64)
    return gain.numpy(), real.numpy(), imag.numpy()


def accum_sha256(out) -> str:
    import hashlib

    return hashlib.sha256(b"".join(np.ascontiguousarray(np.asarray(out[k])).tobytes()
                                   for k in ACCUM_FIELDS)).hexdigest()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("coherent", [True, False])
def test_oracle_accumulate_reproduces_reference_hashes(dtype, coherent):
    gain, real, imag = accum_fixture()
    dt = np.float32 if dtype == "float32" else np.float64
    out = oracle.accumulate_flat(np.ones(len(ACCUM_TX), bool), ACCUM_TX, ACCUM_RX, ACCUM_CID, gain.astype(dt),
                                 real.astype(dt), imag.astype(dt), num_tx=2, num_rx=2, coherent=coherent, dtype=dt)
    assert accum_sha256(out) == ACCUM_SHA256[(dtype, coherent)]


# ----------------------------------------------------------------- host compile

def test_coplanar_groups_golden():
    """Reference test_reflection_multibounce.py:83-117: same surface id, two
    planes -> groups [0, 0, 1], representatives [0, 2]."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    faces = np.array([[0, 1, 2], [1, 3, 2], [4, 5, 6]])
    tri_a = pts[faces[:, 0]]
    n = np.cross(pts[faces[:, 1]] - tri_a, pts[faces[:, 2]] - tri_a)
    g = coplanar_face_groups(tri_a, n, np.full(3, 7))
    assert g["group_count"] == 2
    assert g["face_group_id"].tolist() == [0, 0, 1]
    assert g["representative_faces"].tolist() == [0, 2]


def test_box_edges_are_convex_wedges():
    v, f, _ = workloads.box_mesh(0, 0, 2, 3, 4)
    groups = coplanar_face_groups(v[f[:, 0]], np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]),
                                  np.zeros(12))
    edges = build_edges(v, f, groups["face_group_id"], np.zeros(12, np.int32))
    assert len(edges) == 12
    np.testing.assert_allclose(edges["n"], 1.5, atol=1e-6)
    # t0 is perpendicular to the edge and n0 points away from the second face
    d = np.stack([edges["dx"], edges["dy"], edges["dz"]], 1)
    t0 = np.stack([edges["t0x"], edges["t0y"], edges["t0z"]], 1)
    assert np.abs((d * t0).sum(1)).max() < 1e-6


def test_exhaustive_planning_thresholds():
    assert is_exhaustive(4096, 1) and not is_exhaustive(4097, 1)
    assert is_exhaustive(316, 2) and not is_exhaustive(317, 2)
    s = exhaustive_sequences(3, 3)
    assert s.shape == (3 * 2 * 2, 3)
    assert (s[:, 1:] != s[:, :-1]).all()
    assert s.tolist() == sorted(s.tolist())


def test_tiles_never_straddle_transmitters():
    offsets = np.array([0, 100, 101, 300])
    tiles = make_tiles(offsets)
    for t, base in tiles:
        assert offsets[t] <= base < offsets[t + 1]
    assert sum(min(64, offsets[t + 1] - b) for t, b in tiles) == 300


def test_locality_order_is_a_blocked_permutation():
    # a 32 x 24 grid (row-major, 2 m): every 64 consecutive entries of the
    # Z-order form an 8 x 8 block, and the order covers each cell once
    cols, rows = 32, 24
    pts = np.array([[2.0 * c, 2.0 * r, 1.5] for r in range(rows) for c in range(cols)])
    perm = locality_permutation(pts)
    assert sorted(perm.tolist()) == list(range(rows * cols))
    for k in range(0, rows * cols, 64):
        block = pts[perm[k:k + 64]]
        assert np.ptp(block[:, 0]) == 14.0 and np.ptp(block[:, 1]) == 14.0
    # ties (coincident receivers) keep input order
    twin = locality_permutation(np.zeros((5, 3)))
    assert twin.tolist() == [0, 1, 2, 3, 4]


def test_cell_order_maps_each_transmitter_range_onto_itself():
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 50, (40, 3))
    shared = cell_order_table(pts, np.array([0, 40, 80, 120]), True)
    uneven = cell_order_table(pts[:30], np.array([0, 7, 7, 30]), False)
    for order, bounds in ((shared, [0, 40, 80, 120]), (uneven, [0, 7, 7, 30])):
        assert order.dtype == np.int32
        for t in range(len(bounds) - 1):
            seg = order[bounds[t]:bounds[t + 1]]
            assert sorted(seg.tolist()) == list(range(bounds[t], bounds[t + 1]))
    # shared receivers: the same permutation for every transmitter
    assert (shared[40:80] - 40 == shared[:40]).all()
    assert (uneven[7:] - 7 == locality_permutation(pts[7:30])).all()


def test_urban_canyon_compiles():
    c = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(4, 4)), reference_frequency_hz=28e9)
    assert 760 <= c.face_count <= 780
    assert c.group_count == 64 * 6 + 1
    assert c.materials.count == 3


# ----------------------------------------------------------------- C ABI

def test_library_exports_every_declared_symbol():
    from witwin.channel import _native

    lib = _native.load()
    header = (ROOT / "include" / "wcmi.h").read_text()
    declared = set(re.findall(r"^\s*(?:[\w\*]+\s+)+\**(wc_\w+)\(", header, flags=re.M))
    assert declared, "no declarations parsed"
    for name in sorted(declared):
        assert hasattr(lib, name), f"libwcmi.so does not export {name}"
    assert lib.wc_abi_version() == 1
    assert lib.wc_build_target() == b"gfx950"


def test_value_returning_entry_points():
    from witwin.channel import _native

    assert _native.call("wc_vjp_partial_stride", 3) == 10
    assert _native.call("wc_bvh_node_count", 5) == 4
    assert _native.call("wc_bvh_node_count", 1) == 1
    assert _native.call("wc_abi_version") == 1
    assert _native.call("wc_bvh_workspace_bytes", 100) > 0
    assert _native.call("wc_accumulate_workspace_bytes", 10, 2, 3) > 0


def _sah_tables(compiled):
    from witwin.channel import _native

    f = compiled.face_count
    nodes = np.zeros(_native.call("wc_bvh_node_count", f) * 64, np.uint8)
    sorted_tris = np.zeros(f * 48, np.uint8)
    face_tris = np.zeros(f * 48, np.uint8)
    arrays = [np.ascontiguousarray(compiled.vertices, np.float32), np.ascontiguousarray(compiled.faces, np.int32),
              np.ascontiguousarray(compiled.face_group, np.int32), np.ascontiguousarray(compiled.face_material, np.int32)]
    _native.call("wc_bvh_build_sah", *(a.ctypes.data for a in arrays), f, nodes.ctypes.data, sorted_tris.ctypes.data,
                 face_tris.ctypes.data)
    return nodes, sorted_tris.view(np.float32).reshape(f, 12), face_tris.view(np.float32).reshape(f, 12)


def _wide_children(nodes, node):
    """(lo [k,3], hi [k,3], child [k]) of a 64-byte wide node, decoded from
    its 8-bit grid steps exactly as the tracers decode them."""
    rec = nodes[node * 64:(node + 1) * 64]
    fl = rec.view(np.float32)
    u = rec.view(np.uint32)
    it = rec.view(np.int32)
    exps = int(u[3])
    k = exps >> 24
    step = np.array([np.ldexp(1.0, ((exps >> (8 * a)) & 255) - 127) for a in range(3)], np.float32)
    origin = fl[0:3]
    q_lo = np.array([[(int(u[4 + a]) >> (8 * j)) & 255 for a in range(3)] for j in range(k)], np.float32)
    q_hi = np.array([[(int(u[7 + a]) >> (8 * j)) & 255 for a in range(3)] for j in range(k)], np.float32)
    lo = (q_lo * step + origin).astype(np.float32)
    hi = (q_hi * step + origin).astype(np.float32)
    # the decode is exact: the float64 value equals the float32 one
    np.testing.assert_array_equal(lo.astype(np.float64), q_lo.astype(np.float64) * step + origin.astype(np.float64))
    np.testing.assert_array_equal(hi.astype(np.float64), q_hi.astype(np.float64) * step + origin.astype(np.float64))
    return lo, hi, it[12:12 + k]


def _build_case(which):
    if which == "canyon":
        return compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    if which == "heightfield":
        return compile_scene(workloads.heightfield_scene(seed=4, size=48, boxes=12, shape=(2, 2)),
                             reference_frequency_hz=3.5e9)
    n = 1 if which == "one" else 2
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 0, 1], [4, 0, 1], [3, 1, 1]], np.float32)
    faces = np.array([[0, 1, 2], [3, 4, 5]], np.int64)[:n]
    mesh = Mesh(vertices=torch.as_tensor(verts), faces=torch.as_tensor(faces))
    scene = Scene(structures=[Structure(name="tri", geometry=mesh, material=PhysicalMaterial(eps_r=3.0))])
    return compile_scene(scene, reference_frequency_hz=28e9)


@pytest.mark.parametrize("which", ["canyon", "heightfield", "one", "two"])
def test_sah_bvh_tables_are_a_valid_tree(which):
    """Host SAH tree, collapsed to 64-byte 4-wide nodes with quantized child
    boxes (decoded exactly): every triangle is one leaf,
    every wide node is reached once from the root and holds 1..4 children
    (2..4 unless the scene has one triangle), each child box holds its whole
    subtree, the walk's worst-case stack (the unvisited siblings of a path)
    fits the 96-entry stack, and the triangle rows carry group / face /
    material exactly as the device LBVH writes them."""
    compiled = _build_case(which)
    f = compiled.face_count
    nodes, sorted_tris, face_tris = _sah_tables(compiled)
    tri = compiled.vertices[compiled.faces].astype(np.float32)  # [f, 3, 3]
    np.testing.assert_array_equal(face_tris.reshape(f, 3, 4)[:, :, :3], tri)
    np.testing.assert_array_equal(face_tris[:, 3].view(np.int32), compiled.face_group)
    np.testing.assert_array_equal(face_tris[:, 7].view(np.int32), np.arange(f))
    np.testing.assert_array_equal(face_tris[:, 11].view(np.int32), compiled.face_material)
    order = sorted_tris[:, 7].view(np.int32)
    assert sorted(order.tolist()) == list(range(f))
    np.testing.assert_array_equal(sorted_tris, face_tris[order])
    seen_leaf = np.zeros(f, np.int64)
    seen_node = {}
    worst_stack = 0

    def visit(node, held):
        nonlocal worst_stack
        seen_node[node] = seen_node.get(node, 0) + 1
        lo, hi, child = _wide_children(nodes, node)
        assert 1 <= len(child) <= 4 and (len(child) >= 2 or f == 1)
        internal = int((child >= 0).sum())
        held = held + max(internal - 1, 0)
        worst_stack = max(worst_stack, held)
        leaves = []
        for k, c in enumerate(child):
            sub = [~c] if c < 0 else visit(int(c), held)
            pts = sorted_tris[sub].reshape(-1, 3, 4)[:, :, :3].reshape(-1, 3)
            assert (pts >= lo[k]).all() and (pts <= hi[k]).all()
            leaves += sub
        return leaves

    for k in visit(0, 0):
        seen_leaf[k] += 1
    assert (seen_leaf == 1).all()
    assert all(v == 1 for v in seen_node.values())
    assert len(seen_node) <= max(f - 1, 1)
    assert worst_stack <= 96


def test_wide_node_levels_put_children_before_parents():
    """Refit order of the wide tree: every reachable node once, each internal
    child in a deeper level (refit earlier) than its parent, the root last."""
    from witwin.channel.scene.compiler import wide_node_levels

    for which in ("canyon", "heightfield", "one"):
        compiled = _build_case(which)
        nodes, _, _ = _sah_tables(compiled)
        order, offsets = wide_node_levels(nodes)
        assert offsets[0] == 0 and offsets[-1] == len(order) and (np.diff(offsets) > 0).all()
        assert order[-1] == 0 and len(set(order.tolist())) == len(order)
        level_of = {int(n): i for i in range(len(offsets) - 1) for n in order[offsets[i]:offsets[i + 1]]}
        for n in order:
            _, _, child = _wide_children(nodes, int(n))
            for c in child:
                if c >= 0:
                    assert level_of[int(c)] < level_of[int(n)]


@pytest.mark.parametrize("which", ["canyon", "heightfield", "one", "two"])
def test_wide_tree_top_levels_take_the_first_ids(which):
    """The walk kernels stage wide nodes 0 .. 340 in LDS: they must be the
    tree's first 341 nodes in breadth-first order from the root (children in
    slot order), so every walk's top levels come from the staged block."""
    compiled = _build_case(which)
    nodes, _, _ = _sah_tables(compiled)
    bfs = [0]
    i = 0
    while i < len(bfs) and len(bfs) < 341:
        _, _, child = _wide_children(nodes, bfs[i])
        bfs += [int(c) for c in child if c >= 0][:341 - len(bfs)]
        i += 1
    assert bfs == list(range(len(bfs)))
    reach, i = [0], 0
    while i < len(reach):
        reach += [int(c) for c in _wide_children(nodes, reach[i])[2] if c >= 0]
        i += 1
    assert len(bfs) == min(341, len(reach))


def _slab_np(o, inv, lo, hi, tmin, tmax):
    t0, t1 = (lo - o) * inv, (hi - o) * inv
    tn = max(np.minimum(t0, t1).max(), tmin)
    tf = min(np.maximum(t0, t1).min() * (1.0 + 32.0 * np.finfo(np.float32).eps), tmax)
    return tn <= tf


def test_wide_walk_reaches_every_crossed_triangle():
    """The any-hit walk over the wide tree (the traversal order of the
    tracers, in numpy) reaches every triangle a segment crosses: for random
    segments through the canyon, the set of triangles whose Moller-Trumbore
    test hits equals the set reached through boxes the segment crosses."""
    compiled = _build_case("canyon")
    nodes, sorted_tris, _ = _sah_tables(compiled)
    tri = sorted_tris.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    v0, e1, e2 = tri[:, 0], tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]
    rng = np.random.default_rng(11)
    for _ in range(300):
        a = np.array([rng.uniform(-10, 200), rng.uniform(-10, 200), rng.uniform(0.5, 45)])
        b = np.array([rng.uniform(-10, 200), rng.uniform(-10, 200), rng.uniform(0.5, 45)])
        d = b - a
        with np.errstate(divide="ignore"):
            inv = 1.0 / d
        p = np.cross(d, e2)
        det = (e1 * p).sum(1)
        ok = np.abs(det) > 1e-12
        inv_det = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        tv = a - v0
        u = (tv * p).sum(1) * inv_det
        q = np.cross(tv, e1)
        v = (d * q).sum(1) * inv_det
        t = (e2 * q).sum(1) * inv_det
        crossed = set(np.nonzero(ok & (u >= 1e-6) & (v >= 1e-6) & (u + v <= 1 - 1e-6) & (t > 1e-4) & (t < 1 - 1e-4))[0])
        reached, stack = set(), [0]
        while stack:
            lo, hi, child = _wide_children(nodes, stack.pop())
            for k, c in enumerate(child):
                if _slab_np(a, inv, lo[k], hi[k], 0.0, 1.0):
                    if c < 0:
                        reached.add(int(~c))
                    else:
                        stack.append(int(c))
        assert crossed <= reached


def test_native_errors_raise_without_gpu_work():
    from witwin.channel import _native

    with pytest.raises(RuntimeError, match="count must be non-negative"):
        _native.call("wc_field_free_space", -1, *([None] * 5), 1.0, *([None] * 7), None)


def test_with_vertices_matches_a_fresh_edge_table_and_group_planes():
    from witwin.channel.scene.compiler import build_edge_table, with_vertices

    scene = workloads.urban_canyon_scene(seed=2, shape=(2, 2), spacing=10.0)
    compiled = compile_scene(scene, reference_frequency_hz=28e9)
    rng = np.random.default_rng(5)
    # one jitter per distinct position, so coincident vertices of different structures stay welded
    _, inverse = np.unique(compiled.vertices, axis=0, return_inverse=True)
    jitter = rng.normal(scale=1e-3, size=(int(inverse.max()) + 1, 3))
    v = compiled.vertices.astype(np.float64) + jitter[inverse.reshape(-1)]
    moved = with_vertices(compiled, v)
    edges, ev = build_edge_table(moved.vertices, compiled.faces, compiled.face_group, compiled.face_material,
                                 compiled.edge_policy)
    # a 1 mm jitter changes no edge's status, so a fresh table lists the same edges
    assert np.array_equal(ev, compiled.edge_vertices)
    for name in edges.dtype.names:
        np.testing.assert_array_equal(moved.edges[name], edges[name], err_msg=name)
    # group planes follow the representative faces
    f = compiled.faces[compiled.representative_faces]
    a, b, c = (moved.vertices.astype(np.float64)[f[:, k]] for k in range(3))
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n *= np.where(n[np.arange(len(n)), np.argmax(np.abs(n), axis=1)] < 0, -1.0, 1.0)[:, None]
    np.testing.assert_allclose(moved.group_plane[:, :3], n, atol=1e-12)
    np.testing.assert_allclose(moved.group_plane[:, 3], (n * a).sum(axis=1), atol=1e-9)
    assert moved.device is None and compiled.vertices.dtype == moved.vertices.dtype


def test_edge_vertices_reproduce_edge_records():
    v, f, _ = workloads.box_mesh(0, 0, 2, 3, 4)
    groups = coplanar_face_groups(v[f[:, 0]], np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]),
                                  np.zeros(12))
    from witwin.channel.scene.compiler import build_edge_table

    edges, ev = build_edge_table(v, f, groups["face_group_id"], np.zeros(12, np.int32))
    assert ev.shape == (12, 4) and (ev[:, 3] >= 0).all()
    np.testing.assert_allclose(np.stack([edges["ax"], edges["ay"], edges["az"]], 1), v[ev[:, 0]])
    d = v[ev[:, 1]] - v[ev[:, 0]]
    np.testing.assert_allclose(edges["length"], np.linalg.norm(d, axis=1), rtol=1e-6)
    # a, b, c0 are the vertices of face 0 in its winding order (a cyclic rotation)
    for e in range(12):
        face = list(f[edges["face0"][e]])
        k = face.index(ev[e, 0])
        assert [face[(k + 1) % 3], face[(k + 2) % 3]] == [ev[e, 1], ev[e, 2]]
        assert ev[e, 3] in f[edges["face1"][e]] and ev[e, 3] not in face


def test_vertex_csr_lists_every_contribution_once():
    from witwin.channel.scene.compiler import vertex_csr

    faces = np.array([[0, 1, 2], [2, 1, 3], [3, 4, 0]])
    group_face = np.array([0, 2])
    edge_vertices = np.array([[0, 1, 2, 3], [1, 3, 2, -1]])
    offset, slot = vertex_csr(6, faces, group_face, edge_vertices)
    assert offset.tolist()[-1] == 6 + 7 and offset[5] == offset[6]
    expect = {v: [] for v in range(6)}
    for g, fc in enumerate(group_face):
        for k in range(3):
            expect[int(faces[fc][k])].append(3 * g + k)
    for e in range(2):
        for k in range(4):
            if edge_vertices[e, k] >= 0:
                expect[int(edge_vertices[e, k])].append(6 + 4 * e + k)
    for v in range(6):
        assert slot[offset[v]:offset[v + 1]].tolist() == expect[v]


def test_heightfield_workload_shape():
    scene = workloads.heightfield_scene(shape=(4, 4))
    compiled = compile_scene(scene, reference_frequency_hz=3.5e9)
    assert compiled.face_count == 2 * 127 * 127 + 40 * 12
    assert compiled.materials.count == 4
    z = compiled.vertices[:128 * 128, 2]
    assert 0.0 <= z.min() and abs(z.max() - 20.0) < 1e-4
    # vertical-only edges: the 4 vertical edges of each box
    assert compiled.edge_count == 160 and np.all(np.abs(compiled.edges["dz"]) > 0.7)
    rx = compiled.endpoints.rx_position
    assert np.allclose(rx[:, 2], 21.5)


def _sommerfeld_half_plane(k, rho, phi, phi_p, soft=True):
    """Exact total field of a perfectly conducting half-plane (faces at phi = 0
    and 2 pi) lit by the unit plane wave e^{jk rho cos(phi - phi')} (time
    factor e^{jwt}): u = v(phi - phi') -+ v(phi + phi'),
    v(psi) = e^{jk rho cos psi} e^{j pi/4} / sqrt(pi) int_{-inf}^{a} e^{-j t^2} dt,
    a = sqrt(2 k rho) cos(psi / 2) (Sommerfeld 1896; Balanis, Advanced
    Engineering Electromagnetics, ch. 13)."""
    from scipy.special import fresnel

    def v(psi):
        a = math.sqrt(2.0 * k * rho) * math.cos(psi / 2.0)
        s, c = fresnel(a * math.sqrt(2.0 / math.pi))
        integral = math.sqrt(math.pi / 2.0) * (complex(c, -s) + complex(0.5, -0.5))
        return cmath.exp(1j * k * rho * math.cos(psi)) * cmath.exp(1j * math.pi / 4) / math.sqrt(math.pi) * integral

    return v(phi - phi_p) - v(phi + phi_p) if soft else v(phi - phi_p) + v(phi + phi_p)


@pytest.mark.parametrize("soft", [True, False])
def test_utd_half_plane_matches_sommerfeld(soft):
    """Geometrical optics plus the oracle's UTD diffracted field of a PEC
    half-plane (n = 2, plane-wave incidence, L = rho) against the exact
    Sommerfeld solution, at every observation angle including both shadow
    boundaries. A wrong sign, cot argument or transition-function convention
    in any of the four terms gives an O(1) error."""
    k = 2 * math.pi / 0.1
    rho = 3.0
    phi_p = 1.1
    r = -1.0 if soft else 1.0
    worst = 0.0
    for phi in np.linspace(0.05, 2 * math.pi - 0.05, 181):
        go = 0.0
        if abs(phi - phi_p) < math.pi:
            go += cmath.exp(1j * k * rho * math.cos(phi - phi_p))
        if phi + phi_p < math.pi:
            go += r * cmath.exp(1j * k * rho * math.cos(phi + phi_p))
        ds, dh = oracle.utd(2.0, float(phi), phi_p, 1.0, k, rho)
        d = ds if soft else dh
        total = go + d * cmath.exp(-1j * k * rho) / math.sqrt(rho)
        exact = _sommerfeld_half_plane(k, rho, float(phi), phi_p, soft)
        worst = max(worst, abs(total - exact))
    assert worst < 1e-8, worst


def _finite(n, phi, phi_p, k, L, trunc, gamma):
    ds, _ = oracle.utd_finite(n, phi, phi_p, 1.0, k, L, trunc, gamma)
    return ds


def test_utd_finite_edge_assembly_keeps_go_compensation():
    """E1d assembly (reference docs/dev/audit/utd-continuity-fix-design.md:
    160-205). With the corner blend gamma = 1 the coefficient's jump across
    the incident and both reflection shadow boundaries is the infinite
    wedge's, whatever the truncation factor T (the step-carrying odd part is
    not multiplied by T); with gamma = 0 the jump vanishes; outside the
    locality window the coefficient is T times the infinite wedge's. The
    E1c form (T times the whole coefficient) fails the first check."""
    k, L, n, phi_p = 2 * math.pi / 0.1, 4.0, 1.5, 0.7
    trunc = 1.27 * cmath.exp(1j * math.radians(15.0))
    d = 1e-9
    for boundary in (math.pi + phi_p, math.pi - phi_p, (2 * n - 1) * math.pi - phi_p):
        plain = _finite(n, boundary + d, phi_p, k, L, 1.0, 1.0) - _finite(n, boundary - d, phi_p, k, L, 1.0, 1.0)
        kept = _finite(n, boundary + d, phi_p, k, L, trunc, 1.0) - _finite(n, boundary - d, phi_p, k, L, trunc, 1.0)
        gone = _finite(n, boundary + d, phi_p, k, L, trunc, 0.0) - _finite(n, boundary - d, phi_p, k, L, trunc, 0.0)
        assert abs(plain) > 1e-2
        assert abs(kept - plain) < 1e-6 * abs(plain), (boundary, kept, plain)
        assert abs(gone) < 1e-6 * abs(plain), (boundary, gone)
    far = 2.0  # radians from every boundary: w = exp(-(2 / 0.5)^2 ...) underflows the window
    for phi in (phi_p + far, 3.0 * math.pi / 2.0 + 0.4):
        if min(abs(phi - b) for b in (math.pi + phi_p, math.pi - phi_p, 2 * math.pi - phi_p)) < 1.0:
            continue
        base = _finite(n, phi, phi_p, k, L, 1.0, 1.0)
        assert abs(_finite(n, phi, phi_p, k, L, trunc, 0.3) - trunc * base) < 1e-9 * abs(base)


# reference tests/deterministic/test_diffraction_single_wedge.py:41-66 (edge
# ids and lengths of the PEC wedge at 3 GHz; pure geometry)
WEDGE_EDGE_IDS = (0, 1, 2, 5)
WEDGE_LENGTHS = (3.650281668, 4.744879723, 3.768759727, 4.046976566)


def test_wedge_edge_records_follow_reference_numbering():
    """Public edge ids of the two-structure wedge: per-structure vertex-pair
    records, the second copy of the shared edge (3) merged into 0. The
    vertical-only policy keeps 0 and the slanted outline edges 2 and 5, which
    is the reference's three rows (test_diffraction_single_wedge.py:88-140)."""
    from witwin.channel import workloads
    from witwin.channel.scene.compiler import compile as compile_scene

    c = compile_scene(workloads.wedge_diffraction_scene(), reference_frequency_hz=3.0e9)
    assert list(c.edge_public_id) == [0, 1, 2, 4, 5]
    e = c.edges
    dz = np.abs(np.asarray(e["dz"]))
    assert dz[0] == pytest.approx(1.0) and float(e["n"][0]) == pytest.approx(1.5)  # shared wedge, 270 deg exterior
    assert dz[1] == pytest.approx(0.0, abs=1e-7) and dz[3] == pytest.approx(0.0, abs=1e-7)  # horizontal 1 and 4
    assert dz[2] == pytest.approx(3.0 / math.sqrt(13.0)) and dz[4] == pytest.approx(3.0 / math.sqrt(13.0))
    v = compile_scene(workloads.wedge_diffraction_scene("vertical_only"), reference_frequency_hz=3.0e9)
    assert list(v.edge_public_id) == [0, 2, 5]


def test_oracle_wedge_path_lengths_match_reference_golden():
    """The oracle's diffraction rows of the reference wedge fixture: one row per
    public edge id 0, 1, 2, 5 with the reference path lengths (rtol 1e-5)."""
    from witwin.channel import workloads
    from witwin.channel.scene.compiler import compile as compile_scene

    c = compile_scene(workloads.wedge_diffraction_scene(), reference_frequency_hz=3.0e9)
    orc = oracle.Oracle(c)
    rows = orc.pair_paths(tx=(0.0, -1.0, 0.5), rx=(3.0, 1.0, 0.5), frequency=3.0e9, components=4, max_depth=1,
                          chains=np.zeros((0, 5), np.int32), chain_depth=np.zeros(0, np.int32))
    lengths = {int(c.edge_public_id[p.edge]): p.length for p in rows}
    for eid, expect in zip(WEDGE_EDGE_IDS, WEDGE_LENGTHS):
        assert lengths[eid] == pytest.approx(expect, rel=1e-5)


def test_unsupported_ad_modes_raise_before_any_launch():
    """max_paths truncation with either AD mode raises before any launch (the
    adjoint and the tangents differentiate every path of a cell)."""
    from witwin.channel import workloads
    from witwin.channel.deterministic import Config, solve

    scene = workloads.two_ray_scene(shape=(4, 4))
    for mode in ("vjp", "jvp"):
        with pytest.raises(NotImplementedError, match="max_paths"):
            solve(scene, Config(ad_mode=mode, max_paths=3), reference_frequency_hz=3.5e9)


def test_build_info_names_the_hip_backend():
    from witwin.channel import build_info

    info = build_info()
    assert info["backend"] == "hip" and info["uses_hip_native"]
    assert not info["uses_rayd_native"] and not info["cuda_available"]
