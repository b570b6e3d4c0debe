"""FP64 CPU oracle of the propagation hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package, and only as the checker (or the timed CPU
baseline). The product path (witwin.channel) never imports it.

The C++ library (oracle.cpp, built by oracle/Makefile into liboracle.so)
restates the reference arithmetic in float64 with brute-force segment
visibility. Parity status of each restated function, i.e. what pins it:

* free space ........... closed form lambda/(4 pi d) e^{-jkd} + dipole
                         projection (reference tests/kernels/test_field_transport.py:28-73,
                         tests/support/reference_channel.py:35-80)  -> pinned
* slab Fresnel ......... closed form of test_field_transport.py:13-24 and its
                         known answer :116-159; half-space limit, PEC limit,
                         Brewster null, zero-thickness identity  -> pinned
* reflection sequence .. known answers above + two-wall multibounce golden
                         vectors (tests/deterministic/test_reflection_multibounce.py:126-215:
                         depths, sequences, lengths) + rotation invariance  -> pinned
* Rouard stack ......... vacuum wall = 1, single-interface Fresnel t, product
                         of layers, energy conservation |r|^2 + Re(...)|t|^2 = 1
                         for lossless stacks  -> pinned
* transmission ......... vacuum wall equals line of sight
                         (tests/deterministic/test_transmission.py:88-180)  -> pinned
* layer stack .......... (layer_stack) characteristic-matrix method, independent of
                         the Rouard recursion above and equal to it to 1e-9;
                         normal-incidence impedance formula, vacuum interior
                         phase, lossless energy balance, Brewster TM null,
                         split layer, PEC limit, thin-slab TE closed form
                         (tests/kernels/test_em_layer_stack.py:110-300)  -> pinned
* transmission rows .... (transmission_sequence) vacuum wall equals free space,
                         lossy normal incidence scales by |t_TE|, invalid rows
                         zero (tests/kernels/test_field_transmission_sequence.py:67-172)
                         -> pinned
* Fresnel integrals .... scipy.special.fresnel  -> pinned
* UTD wedge ............ geometrical-optics compensation across the incident
                         and reflection shadow boundaries (total field continuous),
                         PEC half-plane against the exact Sommerfeld solution;
                         the reference's own wedge golden gains
                         (test_diffraction_single_wedge.py:41-85) depend on RayD
                         edge records and blending that are not in the reference
                         tree  -> partial (edge semantics and lossy-wedge
                         coefficients: parity unpinned)
* accumulation ......... slot map and coherent/incoherent finalize of
                         deterministic.cu:829-970 on hand-built rows; the
                         reference-arithmetic restatement (accumulate_flat)
                         reproduces the reference's four frozen SHA-256
                         output hashes (tests/ad/test_deterministic_accum_ad.py:46-129:
                         float32/float64, coherent/incoherent)  -> pinned
* canonical order ...... sort-order golden vector of
                         test_reflection_multibounce.py:25-80  -> pinned
* canonical compact .... (canonical_compact) the sort-order golden vector above,
                         hand-worked identity groups (shortest row kept, first
                         of equal shortest, a longer duplicate dropped, a
                         diffraction row keyed by its edge) and the per-pair
                         cap  -> pinned
* discovery rays ....... restated ray-by-ray; compared with the GPU chain set  -> pinned by equality
* coupled paths ........ (Oracle.coupled_paths) R->D / D->R fields equal the image-theory first-order
                         diffraction over a perfect conductor; every path's length is stationary
                         (Fermat) with its points on their plane and edges  -> pinned; the D->D
                         second-edge spreading factor has no such check  -> parity unpinned
"""

from __future__ import annotations

import cmath
import ctypes
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib: ctypes.CDLL | None = None

MAXD = 5


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


class OracleScene(ctypes.Structure):
    _fields_ = [
        ("face_count", ctypes.c_int), ("tri", ctypes.c_void_p), ("face_group", ctypes.c_void_p),
        ("face_material", ctypes.c_void_p), ("group_count", ctypes.c_int), ("plane", ctypes.c_void_p),
        ("group_off", ctypes.c_void_p), ("group_faces", ctypes.c_void_p), ("mat_eps", ctypes.c_void_p),
        ("mat_sigma", ctypes.c_void_p), ("mat_mu", ctypes.c_void_p), ("mat_gain", ctypes.c_void_p),
        ("mat_thick", ctypes.c_void_p), ("layer_off", ctypes.c_void_p), ("layer_cnt", ctypes.c_void_p),
        ("layer_thick", ctypes.c_void_p), ("layer_eps", ctypes.c_void_p), ("layer_sigma", ctypes.c_void_p),
        ("layer_mu", ctypes.c_void_p), ("edge_count", ctypes.c_int), ("edge_f", ctypes.c_void_p),
        ("edge_i", ctypes.c_void_p),
    ]


class OracleSolve(ctypes.Structure):
    _fields_ = [
        ("tx", ctypes.c_double * 3), ("tx_pol", ctypes.c_double * 3), ("tx_power", ctypes.c_double),
        ("rx", ctypes.c_double * 3), ("rx_pol", ctypes.c_double * 3), ("frequency", ctypes.c_double),
        ("components", ctypes.c_int), ("max_depth", ctypes.c_int), ("chain_count", ctypes.c_int),
        ("chain_groups", ctypes.c_void_p), ("chain_depth", ctypes.c_void_p),
    ]


class OraclePath(ctypes.Structure):
    _fields_ = [
        ("component", ctypes.c_int), ("depth", ctypes.c_int), ("prim", ctypes.c_int * MAXD),
        ("mat", ctypes.c_int * MAXD), ("edge", ctypes.c_int), ("length", ctypes.c_double),
        ("pos", (ctypes.c_double * 3) * MAXD), ("field", ctypes.c_double * 6),
        ("coefficient", ctypes.c_double * 2), ("path_field", ctypes.c_double * 2), ("gain", ctypes.c_double),
        ("arrival", ctypes.c_double * 3), ("departure", ctypes.c_double * 3),
    ]


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        lib.oracle_slab_fresnel.argtypes = [d, d, d, d, d, d, d, vp]
        lib.oracle_stack_t.argtypes = [d, vp, vp, vp, vp, i, d, i, vp]
        lib.oracle_transition.argtypes = [d, vp]
        lib.oracle_fresnel.argtypes = [d, vp]
        lib.oracle_utd.argtypes = [d, d, d, d, d, d, vp, vp]
        lib.oracle_discover_all.argtypes = [ctypes.POINTER(OracleScene), vp, i, i, vp]
        lib.oracle_utd_finite.argtypes = [d, d, d, d, d, d, vp, vp, d, vp]
        lib.oracle_free_space.argtypes = [vp, vp, d, vp, vp, d, vp]
        lib.oracle_reflection_sequence.argtypes = [i, vp, vp, vp, vp, d, vp, vp, vp, vp, vp, vp, vp, d, vp]
        lib.oracle_pair_paths.argtypes = [ctypes.POINTER(OracleScene), ctypes.POINTER(OracleSolve),
                                          ctypes.POINTER(OraclePath), i]
        lib.oracle_pair_paths.restype = i
        lib.oracle_coupled_paths.argtypes = [ctypes.POINTER(OracleScene), ctypes.POINTER(OracleSolve), i,
                                             ctypes.POINTER(OraclePath), i]
        lib.oracle_coupled_paths.restype = i
        lib.oracle_pair_accumulate.argtypes = [ctypes.POINTER(OraclePath), i, i, d, vp]
        lib.oracle_radiomap.argtypes = [ctypes.POINTER(OracleScene), ctypes.POINTER(OracleSolve), vp, i, i, vp,
                                        vp, i, vp, d, vp]
        lib.oracle_discover_ray.argtypes = [ctypes.POINTER(OracleScene), vp, i, i, i, vp]
        lib.oracle_path_record_bytes.restype = i
        if lib.oracle_path_record_bytes() != ctypes.sizeof(OraclePath):
            raise RuntimeError("OraclePath layout mismatch")
        _lib = lib
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


# ----------------------------------------------------------------- unit operators

def slab_fresnel(cos_theta, eps_r, sigma_e, mu_r, gain, thickness, frequency) -> tuple[complex, complex]:
    out = np.zeros(4)
    load().oracle_slab_fresnel(cos_theta, eps_r, sigma_e, mu_r, gain, thickness, frequency, _p(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


def stack_t(cos_theta, thickness, eps_r, sigma_e, mu_r, frequency, pol: int) -> complex:
    th, er, sg, mr = (_c(x, np.float32) for x in (thickness, eps_r, sigma_e, mu_r))
    out = np.zeros(2)
    load().oracle_stack_t(cos_theta, _p(th), _p(er), _p(sg), _p(mr), int(th.shape[0]), frequency, pol, _p(out))
    return complex(out[0], out[1])


def transition(x: float) -> complex:
    out = np.zeros(2)
    load().oracle_transition(float(x), _p(out))
    return complex(out[0], out[1])


def fresnel(x: float) -> tuple[float, float]:
    out = np.zeros(2)
    load().oracle_fresnel(float(x), _p(out))
    return float(out[0]), float(out[1])


def utd(n, phi, phi_p, sin_beta0, k, L, r0s=-1, rns=-1, r0h=1, rnh=1) -> tuple[complex, complex]:
    r = _c([complex(r0s).real, complex(r0s).imag, complex(rns).real, complex(rns).imag,
            complex(r0h).real, complex(r0h).imag, complex(rnh).real, complex(rnh).imag], np.float64)
    out = np.zeros(4)
    load().oracle_utd(n, phi, phi_p, sin_beta0, k, L, _p(r), _p(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


def utd_finite(n, phi, phi_p, sin_beta0, k, L, trunc, gamma, r0s=-1, rns=-1, r0h=1, rnh=1) -> tuple[complex, complex]:
    """Wedge coefficients with the finite-edge truncation factor ``trunc``
    and corner blend ``gamma`` folded into the terms (E1d assembly)."""
    r = _c([complex(r0s).real, complex(r0s).imag, complex(rns).real, complex(rns).imag,
            complex(r0h).real, complex(r0h).imag, complex(rnh).real, complex(rnh).imag], np.float64)
    t = _c([complex(trunc).real, complex(trunc).imag], np.float64)
    out = np.zeros(4)
    load().oracle_utd_finite(n, phi, phi_p, sin_beta0, k, L, _p(r), _p(t), float(gamma), _p(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


def free_space(src, tgt, power, tx_pol, rx_pol, frequency) -> dict:
    out = np.zeros(12)
    a, b, tp, rp = (_c(x, np.float64) for x in (src, tgt, tx_pol, rx_pol))
    load().oracle_free_space(_p(a), _p(b), float(power), _p(tp), _p(rp), float(frequency), _p(out))
    return _row(out)


def reflection_sequence(src, tgt, positions, normals, power, tx_pol, rx_pol, eps, sigma, mu, gain, thickness,
                        frequency) -> dict:
    pos = _c(positions, np.float64).reshape(-1, 3)
    depth = pos.shape[0]
    arrs = [_c(x, np.float64) for x in (src, tgt, pos, normals, tx_pol, rx_pol, eps, sigma, mu, gain, thickness)]
    out = np.zeros(12)
    load().oracle_reflection_sequence(depth, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), float(power),
                                      _p(arrs[4]), _p(arrs[5]), _p(arrs[6]), _p(arrs[7]), _p(arrs[8]), _p(arrs[9]),
                                      _p(arrs[10]), float(frequency), _p(out))
    return _row(out)


def _row(out: np.ndarray) -> dict:
    return {
        "field_vector": out[0:6:2] + 1j * out[1:6:2],
        "coefficient": complex(out[6], out[7]),
        "path_field": complex(out[8], out[9]),
        "path_gain": float(out[10]),
        "path_length": float(out[11]),
    }


# ----------------------------------------------------------------- scene oracle

class Oracle:
    """Oracle view of a compiled scene (host arrays from witwin.channel's compile)."""

    def __init__(self, compiled) -> None:
        self.lib = load()
        m = compiled.materials
        self._keep = {
            "tri": _c(compiled.face_vertices(), np.float32),
            "face_group": _c(compiled.face_group, np.int32),
            "face_material": _c(compiled.face_material, np.int32),
            "plane": _c(compiled.group_plane, np.float64),
            "group_off": _c(compiled.group_face_offset, np.int32),
            "group_faces": _c(compiled.group_faces, np.int32),
            "mat_eps": _c(m.eps_r, np.float32), "mat_sigma": _c(m.sigma_e, np.float32),
            "mat_mu": _c(m.mu_r, np.float32), "mat_gain": _c(m.gain, np.float32),
            "mat_thick": _c(m.thickness_m, np.float32), "layer_off": _c(m.layer_offset, np.int32),
            "layer_cnt": _c(m.layer_count, np.int32), "layer_thick": _c(m.layer_thickness_m, np.float32),
            "layer_eps": _c(m.layer_eps_r, np.float32), "layer_sigma": _c(m.layer_sigma_e, np.float32),
            "layer_mu": _c(m.layer_mu_r, np.float32),
        }
        e = compiled.edges
        fnames = ("ax", "ay", "az", "dx", "dy", "dz", "length", "n", "t0x", "t0y", "t0z", "n0x", "n0y", "n0z")
        inames = ("face0", "face1", "group0", "group1", "material0", "material1")
        self._keep["edge_f"] = _c(np.stack([e[n] for n in fnames], axis=1) if len(e) else np.zeros((0, 14)),
                                  np.float32)
        self._keep["edge_i"] = _c(np.stack([e[n] for n in inames], axis=1) if len(e) else np.zeros((0, 6)),
                                  np.int32)
        k = self._keep
        self.scene = OracleScene(
            face_count=compiled.face_count, tri=_p(k["tri"]), face_group=_p(k["face_group"]),
            face_material=_p(k["face_material"]), group_count=compiled.group_count, plane=_p(k["plane"]),
            group_off=_p(k["group_off"]), group_faces=_p(k["group_faces"]), mat_eps=_p(k["mat_eps"]),
            mat_sigma=_p(k["mat_sigma"]), mat_mu=_p(k["mat_mu"]), mat_gain=_p(k["mat_gain"]),
            mat_thick=_p(k["mat_thick"]), layer_off=_p(k["layer_off"]), layer_cnt=_p(k["layer_cnt"]),
            layer_thick=_p(k["layer_thick"]), layer_eps=_p(k["layer_eps"]), layer_sigma=_p(k["layer_sigma"]),
            layer_mu=_p(k["layer_mu"]), edge_count=compiled.edge_count, edge_f=_p(k["edge_f"]),
            edge_i=_p(k["edge_i"]),
        )

    def _solve(self, tx, tx_pol, tx_power, rx_pol, frequency, components, max_depth, chains, chain_depth):
        q = OracleSolve()
        q.tx[:] = [float(x) for x in tx]
        q.tx_pol[:] = [float(x) for x in tx_pol]
        q.tx_power = float(tx_power)
        q.rx_pol[:] = [float(x) for x in rx_pol]
        q.frequency = float(frequency)
        q.components = int(components)
        q.max_depth = int(max_depth)
        groups = _c(chains, np.int32).reshape(-1, MAXD)
        depth = _c(chain_depth, np.int32).reshape(-1)
        q.chain_count = int(groups.shape[0])
        q.chain_groups = _p(groups) if groups.size else None
        q.chain_depth = _p(depth) if depth.size else None
        return q, (groups, depth)

    def radiomap(self, *, tx, rx, tx_pol=(0, 0, 1), tx_power=1.0, rx_pol=(0, 0, 1), frequency, components,
                 max_depth, chains, chain_depth, coherent=True, paths_cap=0, perturb=0.0):
        """Per-receiver accumulate records [n, 22] (+ paths and ambiguity flags)."""
        q, keep = self._solve(tx, tx_pol, tx_power, rx_pol, frequency, components, max_depth, chains, chain_depth)
        rx = _c(rx, np.float64).reshape(-1, 3)
        n = rx.shape[0]
        out = np.zeros((n, 22))
        counts = np.zeros(n, np.int32)
        amb = np.zeros(n, np.int32) if perturb > 0 else None
        paths = (OraclePath * (n * paths_cap))() if paths_cap > 0 else None
        self.lib.oracle_radiomap(ctypes.byref(self.scene), ctypes.byref(q), _p(rx), n, 1 if coherent else 0,
                                 _p(out), ctypes.cast(paths, ctypes.c_void_p) if paths is not None else None,
                                 paths_cap, _p(counts), float(perturb), _p(amb) if amb is not None else None)
        result = {"record": out, "counts": counts, "ambiguous": amb}
        if paths is not None:
            result["paths"] = [[paths[i * paths_cap + j] for j in range(min(int(counts[i]), paths_cap))]
                               for i in range(n)]
        return result

    def pair_paths(self, *, tx, rx, tx_pol=(0, 0, 1), tx_power=1.0, rx_pol=(0, 0, 1), frequency, components,
                   max_depth, chains, chain_depth) -> list[OraclePath]:
        q, keep = self._solve(tx, tx_pol, tx_power, rx_pol, frequency, components, max_depth, chains, chain_depth)
        q.rx[:] = [float(x) for x in rx]
        cap = 4096
        buf = (OraclePath * cap)()
        count = self.lib.oracle_pair_paths(ctypes.byref(self.scene), ctypes.byref(q), buf, cap)
        return [buf[i] for i in range(min(count, cap))]

    def coupled_paths(self, *, tx, rx, tx_pol=(0, 0, 1), tx_power=1.0, rx_pol=(0, 0, 1), frequency,
                      kinds: int = 7) -> list[OraclePath]:
        """Coupled paths of one pair in candidate order (components 3 R->D, 4
        D->R, 7 D->D; oracle.cpp oracle_coupled_paths)."""
        q, keep = self._solve(tx, tx_pol, tx_power, rx_pol, frequency, 0, 0, np.zeros((0, MAXD)), np.zeros(0))
        q.rx[:] = [float(x) for x in rx]
        cap = 65536
        buf = (OraclePath * cap)()
        count = self.lib.oracle_coupled_paths(ctypes.byref(self.scene), ctypes.byref(q), int(kinds), buf, cap)
        return [buf[i] for i in range(min(count, cap))]

    def discover_all(self, tx, ray_count: int, max_depth: int) -> np.ndarray:
        """[ray_count, max_depth] group hit per bounce of every discovery ray."""
        t = _c(tx, np.float64)
        out = np.full((ray_count, max_depth), -1, np.int32)
        self.lib.oracle_discover_all(ctypes.byref(self.scene), _p(t), int(ray_count), int(max_depth), _p(out))
        return out

    def discover(self, tx, ray_count: int, max_depth: int, rays: np.ndarray | None = None) -> np.ndarray:
        t = _c(tx, np.float64)
        idx = np.arange(ray_count) if rays is None else np.asarray(rays)
        out = np.zeros((idx.shape[0], max_depth), np.int32)
        row = np.zeros(max_depth, np.int32)
        for j, i in enumerate(idx):
            self.lib.oracle_discover_ray(ctypes.byref(self.scene), _p(t), int(i), int(ray_count), int(max_depth),
                                         _p(row))
            out[j] = row
        return out


# ----------------------------------------------------------------- canonical compact

def _event(component: int, depth: int, slot: int, width: int) -> int:
    """Event type of a row's slot (reference evaluated_paths.cu:128-153)."""
    if component == 1 and slot < depth:
        return 1
    if component == 2 and depth > 0 and slot == 0:
        return 2
    if component == 5 and slot < depth:
        return 4
    if component == 6 and depth > 0 and slot == 0:
        return 8
    if width >= 2 and depth >= 2 and slot < 2:
        if component == 3:
            return 1 if slot == 0 else 2
        if component == 4:
            return 2 if slot == 0 else 1
        if component == 7:
            return 2
    return 0


def _identity(rows: dict, i: int, width: int) -> tuple:
    c, d = int(rows["component_id"][i]), int(rows["depth"][i])
    out = []
    for j in range(width):
        t = _event(c, d, j, width)
        obj = 0
        if t:
            obj = int(rows["edge_id"][i]) if (c == 2 and j == 0) else int(rows["primitive_sequence"][i, j])
        out.append((t, obj))
    return tuple(out)


def canonical_compact(rows: dict, pair_offset: np.ndarray, cap: int | None = None) -> dict:
    """Canonical compact of emitted rows, pair by pair (reference
    native/channel/kernels/evaluated_paths.cu): stable sort on (depth,
    component, primitive sequence, primitive, edge) (:654-700); identity groups
    (per slot: event type and object id, :126-310; runs of the sorted pair,
    :216-320) keep their first strictly shortest row, none when a length is
    NaN; at most ``cap`` identity winners per pair (:355-384).
    rows: numpy arrays (``primitive_sequence`` [K, width]). Returns the sorted
    permutation, keep flags in sorted order, the kept rows' offsets per pair and
    their source row indices."""
    width = rows["primitive_sequence"].shape[1]
    pairs = len(pair_offset) - 1
    perm = np.zeros(int(pair_offset[-1]), np.int64)
    keep = np.zeros(int(pair_offset[-1]), np.uint8)
    kept_offset = np.zeros(pairs + 1, np.int64)
    source = []
    dropped = 0
    for p in range(pairs):
        b, e = int(pair_offset[p]), int(pair_offset[p + 1])
        key = lambda i: (int(rows["depth"][i]), int(rows["component_id"][i]),  # noqa: E731
                         *(int(x) for x in rows["primitive_sequence"][i]), int(rows["primitive_id"][i]),
                         int(rows["edge_id"][i]))
        order = sorted(range(b, e), key=key)  # Python's sort is stable
        perm[b:e] = order
        s = 0
        winners = []
        while s < len(order):
            ident = _identity(rows, order[s], width)
            t = s + 1
            while t < len(order) and _identity(rows, order[t], width) == ident:
                t += 1
            lengths = [float(rows["path_length_m"][order[x]]) for x in range(s, t)]
            if not any(np.isnan(lengths)):  # a NaN length leaves its group without a winner
                win = s + lengths.index(min(lengths))
                keep[b + win] = 1
                winners.append(order[win])
            s = t
        kept = winners if cap is None else winners[:cap]
        dropped += len(winners) - len(kept)
        source.extend(kept)
        kept_offset[p + 1] = kept_offset[p] + len(kept)
    return {"perm": perm, "keep": keep, "kept_offset": kept_offset, "source_row": np.asarray(source, np.int64),
            "dropped": dropped}


# ----------------------------------------------------------------- flat accumulation

def _fma_round(a: float, b: float, c: float, dtype) -> float:
    """a*b + c rounded once to ``dtype`` (exact rational intermediate)."""
    from fractions import Fraction

    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    if dtype == np.float64:
        return float(exact)
    # round-to-nearest-even to float32 from the exact value: float64 first can
    # double-round, so pick the nearer of the two float32 neighbours exactly
    f = np.float32(float(exact))
    lo, hi = (np.nextafter(f, np.float32(-np.inf)), f) if Fraction(float(f)) > exact else \
        (f, np.nextafter(f, np.float32(np.inf)))
    dl, dh = exact - Fraction(float(lo)), Fraction(float(hi)) - exact
    if dl < dh:
        return lo
    if dh < dl:
        return hi
    return lo if (lo.view(np.int32) & 1) == 0 else hi


def accumulate_flat(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, *, num_tx: int,
                    num_rx: int, coherent: bool, dtype=np.float32) -> dict:
    """The reference's flat accumulation in ``dtype`` arithmetic (reference
    native/channel/kernels/deterministic.cu:829-970): rows into six slots per
    cell (0/1/2 -> 0/1/2, 5 -> 3, 6 -> 4 scattering, 3/4/7 -> 5), per-slot
    sums in row order, coherent |z|^2 as one fused multiply-add (what the
    reference's compiler emits), scattering a power-domain slot."""
    dt = np.float32 if dtype == np.float32 else np.float64

    def slot(c):
        if 0 <= c < 3:
            return c
        if c in (3, 4, 7):
            return 5
        return {5: 3, 6: 4}.get(int(c), -1)

    cells = num_tx * num_rx
    cp, cr, ci = (np.zeros((6, cells), dt) for _ in range(3))
    for i in range(len(tx_id)):
        s, t, r = slot(int(component_id[i])), int(tx_id[i]), int(rx_id[i])
        if not valid[i] or s < 0 or t < 0 or r < 0 or t >= num_tx or r >= num_rx:
            continue
        c = t * num_rx + r
        cp[s, c] = dt(cp[s, c] + dt(path_gain[i]))
        cr[s, c] = dt(cr[s, c] + dt(field_real[i]))
        ci[s, c] = dt(ci[s, c] + dt(field_imag[i]))
    pt, fr, fi = (np.zeros(cells, dt) for _ in range(3))
    for c in range(cells):
        rs = is_ = ps = sc = dt(0)
        for s in range(6):
            a, b = cr[s, c], ci[s, c]
            if coherent:
                if s == 4:
                    sc = dt(sc + cp[s, c])
                    continue
                cp[s, c] = _fma_round(a, a, dt(b * b), dt)
            else:
                ps = dt(ps + cp[s, c])
            rs, is_ = dt(rs + a), dt(is_ + b)
        if coherent:
            fr[c], fi[c] = rs, is_
            pt[c] = dt(_fma_round(rs, rs, dt(is_ * is_), dt) + sc)
        else:
            pt[c], fr[c], fi[c] = ps, dt(np.sqrt(max(ps, dt(0)))), dt(0)
    shape2, shape3 = (num_tx, num_rx), (6, num_tx, num_rx)
    return {"power_total": pt.reshape(shape2), "field_total_real": fr.reshape(shape2),
            "field_total_imag": fi.reshape(shape2), "component_power": cp.reshape(shape3),
            "component_field_real": cr.reshape(shape3), "component_field_imag": ci.reshape(shape3)}


# ----------------------------------------------------------------- layer stacks and transmission rows

_EPS0 = 8.8541878128e-12
_MU0 = 1.25663706212e-6
_C0 = 299792458.0


def _passive_sqrt(z: complex) -> complex:
    """Square root on the passive branch (imaginary part <= 0)."""
    m = abs(z)
    return complex(np.sqrt(max(0.5 * (m + z.real), 0.0)), -np.sqrt(max(0.5 * (m - z.real), 0.0)))


def layer_stack(layers, cos_theta: float, frequency: float, pol: int) -> tuple[complex, complex]:
    """(r, t) of a layered wall between vacuum half-spaces by the
    characteristic-matrix method (Born & Wolf, Principles of Optics, 1.6):
    each layer contributes [[cos d, j sin d / Y], [j Y sin d, cos d]] with
    d = kz t, Y the tangential admittance (TE kz / (omega mu), TM
    omega eps / kz), eps = eps0 eps_r - j sigma / omega; r = (Y0 B - C) /
    (Y0 B + C), t = 2 Y0 / (Y0 B + C) with [B, C] = M [1, Y0]. ``layers`` is
    a sequence of (thickness, eps_r, sigma_e, mu_r) entry -> exit; pol 0 = TE,
    1 = TM. An independent formulation of what the kernels compute by the
    backward Rouard recursion (reference kernels/materials.py:151
    em_layer_stack_eval; its test oracle tests/kernels/test_em_layer_stack.py:66-107)."""
    omega = 2.0 * np.pi * frequency
    ct = min(max(abs(cos_theta), 1e-6), 1.0)
    k_par = omega / _C0 * np.sqrt(max(0.0, 1.0 - ct * ct))

    def medium(eps_r, sigma, mu_r):
        eps = _EPS0 * max(eps_r, 1e-6) - 1j * max(sigma, 0.0) / omega
        mu = _MU0 * max(mu_r, 1e-6)
        k = omega * _passive_sqrt(eps * mu)
        kz = _passive_sqrt(k * k - k_par * k_par)
        return kz, (kz / (omega * mu) if pol == 0 else omega * eps / kz)

    _, y0 = medium(1.0, 0.0, 1.0)
    m = np.eye(2, dtype=complex)
    scale = 0.0  # the product is kept as m e^{scale}: each layer factor is divided by e^{|Im d|}
    for thickness, eps_r, sigma, mu_r in layers:
        kz, y = medium(eps_r, sigma, mu_r)
        d = kz * max(thickness, 0.0)
        g = abs(d.imag)
        ep, em = cmath.exp(1j * d - g), cmath.exp(-1j * d - g)
        cos_d, sin_d = 0.5 * (ep + em), -0.5j * (ep - em)
        m = m @ np.array([[cos_d, 1j * sin_d / y], [1j * y * sin_d, cos_d]])
        scale += g
    b = m[0, 0] + m[0, 1] * y0
    c = m[1, 0] + m[1, 1] * y0
    return complex((y0 * b - c) / (y0 * b + c)), complex(2.0 * y0 / (y0 * b + c) * np.exp(-scale))


def _unit(v, fallback):
    n = np.linalg.norm(v)
    return v / n if n > 1e-6 else np.asarray(fallback, np.float64)


def _perp_basis(d, preferred):
    """Unit vector perpendicular to d, from ``preferred`` or a world axis."""
    p = preferred - d * np.dot(d, preferred)
    if np.linalg.norm(p) > 1e-6:
        return p / np.linalg.norm(p)
    axis = np.array([1.0, 0.0, 0.0]) if abs(d[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    return _unit(axis - d * np.dot(d, axis), [0, 0, 1])


def transmission_sequence(path_valid, src, tgt, normals, material_id, interaction_valid, tx_power, tx_pol, rx_pol,
                          materials, frequency) -> dict:
    """Straight transmission rows in float64 (reference
    tests/ad/_reference_fields.py:269-336, the restatement of the reference's
    transmission_sequence_kernel): per valid interaction the wall's s/p basis,
    t_TE / t_TM of :func:`layer_stack`, carrier phase over
    L - sum(thickness) cos theta, amplitude 1 / (2 k L). ``materials`` is a
    list of layer lists. Invalid rows, and rows whose valid interaction names
    no material, give zeros."""
    n = len(path_valid)
    out = {"field_vector": np.zeros((n, 3), complex), "coefficient": np.zeros(n, complex),
           "path_field": np.zeros(n, complex), "path_gain": np.zeros(n), "path_length_m": np.zeros(n),
           "delay_s": np.zeros(n), "direction": np.zeros((n, 3))}
    k = 2.0 * np.pi * frequency / _C0
    ez = np.array([0.0, 0.0, 1.0])
    for i in range(n):
        depth = len(interaction_valid[i])
        if not path_valid[i] or any(interaction_valid[i][w] and not 0 <= material_id[i][w] < len(materials)
                                    for w in range(depth)):
            continue
        a, b = np.asarray(src[i], np.float64), np.asarray(tgt[i], np.float64)
        length = float(np.linalg.norm(b - a))
        d = _unit(b - a, ez)
        tp = np.asarray(tx_pol[i], np.float64)
        value = (tp - d * np.dot(d, tp)).astype(complex)
        carrier = length
        for w in range(depth):
            if not interaction_valid[i][w]:
                continue
            layers = materials[material_id[i][w]]
            nrm = _unit(np.asarray(normals[i][w], np.float64), ez)
            if np.dot(d, nrm) > 0.0:
                nrm = -nrm
            ct = min(max(abs(float(np.dot(d, nrm))), 1e-6), 1.0)
            _, t_te = layer_stack(layers, ct, frequency, 0)
            _, t_tm = layer_stack(layers, ct, frequency, 1)
            s = np.cross(nrm, d)
            s = s / np.linalg.norm(s) if np.linalg.norm(s) > 1e-6 else _perp_basis(d, nrm)
            p = np.cross(s, d)
            p = p / np.linalg.norm(p) if np.linalg.norm(p) > 1e-6 else _perp_basis(d, s)
            value = s * (t_te * np.dot(value, s)) + p * (t_tm * np.dot(value, p))
            carrier -= sum(max(layer[0], 0.0) for layer in layers) * ct
        field = value * (np.exp(-1j * k * carrier) / (2.0 * k * max(length, 1e-10)))
        rp = np.asarray(rx_pol[i], np.float64)
        coef = complex(np.dot(field, rp - d * np.dot(d, rp)))
        pf = coef * np.sqrt(max(float(tx_power[i]), 0.0))
        out["field_vector"][i] = field
        out["coefficient"][i] = coef
        out["path_field"][i] = pf
        out["path_gain"][i] = abs(pf) ** 2
        out["path_length_m"][i] = length
        out["delay_s"][i] = length / _C0
        out["direction"][i] = d
    return out
