"""GPU tests of the fixed-topology VJP against central finite differences of
the FP64 oracle (the reference's AD check: tests/ad/_fd.py:20-62 with the
relative frequency step 1e-4 of tests/ad/_tolerances.py:29).

Loss: L = sum_c w_c P_c over the cells whose topology is stable under the
oracle's perturbation probe (w_c seeded uniform). Gate: every gradient
component within 1e-2 relative of the FP64 central difference (float32
tangents vs float64 differences).
"""

from __future__ import annotations

import copy
import sys

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import run_gpu, tx_chains
from witwin.channel import workloads
from witwin.channel.kernels.deterministic import radiomap_vjp_outputs

pytestmark = pytest.mark.gpu
REL_TOL = 1.0e-2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_power(compiled, run, rx_idx, *, eps=None, sigma=None, frequency=None, tx=None):
    c = copy.copy(compiled)
    m = copy.deepcopy(compiled.materials)
    if eps is not None:
        m.eps_r = m.eps_r.copy()
        m.layer_eps_r = m.layer_eps_r.copy()
        for i, v in enumerate(eps):
            m.eps_r[i] = np.float32(v)
            m.layer_eps_r[m.layer_offset[i]] = np.float32(v)
    if sigma is not None:
        m.sigma_e = m.sigma_e.copy()
        m.layer_sigma_e = m.layer_sigma_e.copy()
        for i, v in enumerate(sigma):
            m.sigma_e[i] = np.float32(v)
            m.layer_sigma_e[m.layer_offset[i]] = np.float32(v)
    c.materials = m
    ep = compiled.endpoints
    groups, depth = tx_chains(run, 0)
    orc = oracle.Oracle(c)
    res = orc.radiomap(tx=ep.tx_position[0] if tx is None else tx, rx=ep.rx_position[rx_idx],
                       frequency=run.frequency if frequency is None else frequency, components=run.bits,
                       max_depth=run.depth, chains=groups, chain_depth=depth)
    return res["record"][:, 0]


def _check(name, got, ref):
    # the loss is normalized to L = 1, so 1e-3 is an absolute floor for near-zero derivatives
    scale = max(abs(ref), 1e-3)
    rel = abs(got - ref) / scale
    _say(f"  d/d{name}: vjp {got:.6e} fd {ref:.6e} rel {rel:.2e}")
    return rel


def _vjp_case(device, scene, components, depth, frequency, h_eps=1e-3, h_sig_rel=1e-3, h_pos=1e-4):
    run = run_gpu(scene, components, depth, frequency, device, topology_capacity=64)
    compiled = run.compiled
    cells = run.inputs.total_cells
    idx = np.arange(cells)
    ep = compiled.endpoints
    groups, cdepth = tx_chains(run, 0)
    base = oracle.Oracle(compiled).radiomap(tx=ep.tx_position[0], rx=ep.rx_position, frequency=frequency,
                                            components=run.bits, max_depth=depth, chains=groups,
                                            chain_depth=cdepth, perturb=2e-4)
    stable = ~base["ambiguous"].astype(bool)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, cells) * stable
    # normalize so every term is O(1)
    p0 = base["record"][:, 0]
    w = w / max(float((w * p0).sum()), 1e-300)
    m = compiled.materials
    eps = torch.tensor(m.eps_r, device=device, requires_grad=True)
    sig = torch.tensor(m.sigma_e, device=device, requires_grad=True)
    freq = torch.tensor(frequency, dtype=torch.float32, device=device, requires_grad=True)
    txp = torch.tensor(ep.tx_position, device=device, requires_grad=True)
    outs, state = radiomap_vjp_outputs(run.dscene, run.inputs, eps_r=eps, sigma_e=sig, frequency=freq,
                                       tx_position=txp, material_count=m.count)
    power = outs[0]
    assert not bool((state.out.failure != 0).any()), "topology capacity exceeded"
    loss = (power * torch.as_tensor(w, dtype=torch.float32, device=device)).sum()
    loss.backward()
    torch.cuda.synchronize()

    def L(**kw):
        return float((w * _oracle_power(compiled, run, idx, **kw)).sum())

    worst = 0.0
    used = [i for i in range(m.count) if (compiled.face_material == i).any()]
    for i in used:
        e_hi, e_lo = m.eps_r.astype(np.float64).copy(), m.eps_r.astype(np.float64).copy()
        e_hi[i] += h_eps
        e_lo[i] -= h_eps
        fd = (L(eps=e_hi) - L(eps=e_lo)) / (2 * h_eps)
        worst = max(worst, _check(f"eps_r[{i}]", float(eps.grad[i]), fd))
        h = max(float(m.sigma_e[i]) * h_sig_rel, 1e-6)
        s_hi, s_lo = m.sigma_e.astype(np.float64).copy(), m.sigma_e.astype(np.float64).copy()
        s_hi[i] += h
        s_lo[i] -= h
        fd = (L(sigma=s_hi) - L(sigma=s_lo)) / (2 * h)
        worst = max(worst, _check(f"sigma[{i}]", float(sig.grad[i]), fd))
    # relative step 1e-7: at 28 GHz and ~300 m paths the phase k*L moves by ~0.02 rad, keeping the
    # central-difference truncation error ~1e-4 (the reference's 1e-4 step is sized for tiny scenes)
    hf = frequency * 1e-7
    fd = (L(frequency=frequency + hf) - L(frequency=frequency - hf)) / (2 * hf)
    # compared as d/d(ln f) = f dL/df so the absolute floor does not hide the check
    worst = max(worst, _check("ln f", float(freq.grad) * frequency, fd * frequency))
    for a in range(3):
        t_hi = ep.tx_position[0].astype(np.float64).copy()
        t_lo = t_hi.copy()
        t_hi[a] += h_pos
        t_lo[a] -= h_pos
        fd = (L(tx=t_hi) - L(tx=t_lo)) / (2 * h_pos)
        worst = max(worst, _check(f"tx[{a}]", float(txp.grad[0, a]), fd))
    return worst


def test_vjp_two_ray_matches_fd(device):
    _say("case: VJP two-ray 32x32, 3.5 GHz, los + ground reflection")
    scene = workloads.two_ray_scene(shape=(32, 32), spacing=2.0)
    worst = _vjp_case(device, scene, {"los", "reflection"}, 1, 3.5e9)
    assert worst <= REL_TOL


def test_vjp_urban_canyon_matches_fd(device):
    _say("case: VJP urban canyon seed 2, 24x24 at 15 m, depth 2, los+refl+diff, 28 GHz")
    scene = workloads.urban_canyon_scene(seed=2, shape=(24, 24), spacing=15.0)
    worst = _vjp_case(device, scene, {"los", "reflection", "diffraction"}, 2, 28e9)
    assert worst <= REL_TOL


def test_vjp_transmission_walls_match_fd(device):
    _say("case: VJP thin-sheet transmission through two layered walls + ground, 3 GHz")
    from witwin.core import MaterialLayer, Mesh, PhysicalMaterial, Scene, Structure

    def wall(x, mat, sid):
        v = np.array([[x, -6, 0], [x, 6, 0], [x, -6, 8], [x, 6, 8]], np.float32)
        return Structure(Mesh(torch.from_numpy(v), torch.tensor([[0, 1, 2], [1, 3, 2]])), mat, surface_id=sid)

    brick = PhysicalMaterial(layers=(MaterialLayer(0.12, 4.4, 0.02), MaterialLayer(0.05, 2.2, 0.001)))
    glass = PhysicalMaterial(eps_r=6.27, sigma_e=0.0043, thickness_m=0.02)
    ground = PhysicalMaterial(eps_r=5.31, sigma_e=0.0326, thickness_m=5.0)
    scene = Scene(structures=[workloads.ground_structure(-50, -50, 50, 50, ground), wall(3.0, brick, 1),
                              wall(6.5, glass, 2)],
                  endpoints=[workloads.transmitter((0.0, 0.3, 2.0)),
                             workloads.grid((8.0, -4.0, 1.5), (8, 8), 1.0)])
    worst = _vjp_case(device, scene, {"los", "reflection", "transmission"}, 2, 3.0e9)
    assert worst <= REL_TOL


def test_solve_ad_mode_vjp_returns_leaf_gradients(device):
    _say("case: deterministic.solve(ad_mode='vjp') with leaf tensors, mean dB loss")
    from witwin.channel.deterministic import Config, solve
    from witwin.core import PhysicalMaterial, Scene

    eps = torch.tensor(5.31, requires_grad=True)
    ground = PhysicalMaterial(eps_r=eps, sigma_e=0.0326, thickness_m=5.0)
    tx = workloads.transmitter((0.0, 0.0, 10.0))
    tx.position.requires_grad_(True)
    scene = Scene(structures=[workloads.ground_structure(-1e4, -1e4, 1e4, 1e4, ground)],
                  endpoints=[tx, workloads.grid((-15.5, -15.5, 1.5), (32, 32), 1.0)])
    freq = torch.tensor(3.5e9, requires_grad=True)
    result = solve(scene, Config(components={"los", "reflection"}, max_depth=1, ad_mode="vjp"),
                   reference_frequency_hz=freq)
    p = result.path_gain
    loss = (10.0 * torch.log10(p)).mean()
    loss.backward()
    assert eps.grad is not None and torch.isfinite(eps.grad)
    assert tx.position.grad is not None and torch.isfinite(tx.position.grad).all()
    assert freq.grad is not None and torch.isfinite(freq.grad)
    _say(f"  grads: eps {float(eps.grad):.4e} tx {tx.position.grad.tolist()} f {float(freq.grad):.4e}")


# ------------------------------------------------------------------ mesh vertices

def _signatures(res) -> list[tuple]:
    return [tuple((p.component, p.depth, tuple(p.prim[:p.depth]), p.edge) for p in cell) for cell in res["paths"]]


def _vertex_loss_terms(compiled, run, vertices, cap=128):
    """Oracle records and path signatures of the scene with moved vertices
    (same face groups and edge list: scene.compiler.with_vertices)."""
    from witwin.channel.scene.compiler import with_vertices

    moved = with_vertices(compiled, vertices)
    ep = compiled.endpoints
    groups, depth = tx_chains(run, 0)
    res = oracle.Oracle(moved).radiomap(tx=ep.tx_position[0], rx=ep.rx_position, frequency=run.frequency,
                                        components=run.bits, max_depth=run.depth, chains=groups, chain_depth=depth,
                                        paths_cap=cap)
    return res["record"][:, 0], _signatures(res)


def _vertex_case(device, scene, components, depth, frequency, picks=4, h=2e-3):
    """dL/d(vertices) of the fixed-topology VJP against central differences of
    the FP64 oracle with one vertex coordinate moved by +-h (Richardson
    extrapolated from steps h and 2h; the oracle's float32 vertex and edge
    tables make the effective step the difference of the rounded
    coordinates). Cells whose oracle topology changes under any probe carry
    no weight."""
    run = run_gpu(scene, components, depth, frequency, device, topology_capacity=64)
    compiled = run.compiled
    cells = run.inputs.total_cells
    ep = compiled.endpoints
    groups, cdepth = tx_chains(run, 0)
    base = oracle.Oracle(compiled).radiomap(tx=ep.tx_position[0], rx=ep.rx_position, frequency=frequency,
                                            components=run.bits, max_depth=depth, chains=groups, chain_depth=cdepth,
                                            perturb=2e-4, paths_cap=128)
    sig0 = _signatures(base)
    stable = ~base["ambiguous"].astype(bool)
    rng = np.random.default_rng(1)
    w_raw = rng.uniform(0.5, 1.5, cells)
    p0 = base["record"][:, 0]
    m = compiled.materials

    def vjp(weights):
        verts = torch.tensor(compiled.vertices, device=device, requires_grad=True)
        outs, state = radiomap_vjp_outputs(run.dscene, run.inputs, eps_r=torch.tensor(m.eps_r, device=device),
                                           sigma_e=torch.tensor(m.sigma_e, device=device),
                                           frequency=torch.tensor(frequency, dtype=torch.float32, device=device),
                                           tx_position=torch.tensor(ep.tx_position, device=device),
                                           material_count=m.count, vertices=verts)
        assert not bool((state.out.failure != 0).any()), "topology capacity exceeded"
        loss = (outs[0] * torch.as_tensor(weights, dtype=torch.float32, device=device)).sum()
        (grad,) = torch.autograd.grad(loss, verts)
        torch.cuda.synchronize()
        return grad.cpu().numpy().astype(np.float64)

    w = w_raw * stable
    w = w / max(float((w * p0).sum()), 1e-300)
    first = vjp(w)
    # probe the vertices with the largest gradients plus one seeded other vertex of the mesh
    mag = np.abs(first).sum(axis=1)
    chosen = list(np.argsort(-mag)[:picks])
    others = [int(i) for i in rng.permutation(len(mag)) if mag[i] > 0 and int(i) not in chosen]
    chosen += others[:1]
    v64 = compiled.vertices.astype(np.float64)
    probes = {}
    for vi in chosen:
        for a in range(3):
            for step in (h, 2 * h):
                for sgn in (1, -1):
                    v = v64.copy()
                    v[vi, a] += sgn * step
                    p, sig = _vertex_loss_terms(compiled, run, v)
                    probes[(vi, a, step, sgn)] = (p, np.float32(v[vi, a]))
                    stable &= np.array([s == s0 for s, s0 in zip(sig, sig0)])
    w = w_raw * stable
    w = w / max(float((w * p0).sum()), 1e-300)
    _say(f"  stable cells {int(stable.sum())} / {cells}; vertices {chosen}")
    grad = vjp(w)
    worst = 0.0
    for vi in chosen:
        for a in range(3):
            fd = []
            for step in (h, 2 * h):
                (p_hi, x_hi), (p_lo, x_lo) = probes[(vi, a, step, 1)], probes[(vi, a, step, -1)]
                fd.append(float((w * (p_hi - p_lo)).sum()) / (float(x_hi) - float(x_lo)))
            ref = (4.0 * fd[0] - fd[1]) / 3.0
            worst = max(worst, _check(f"v[{vi}][{'xyz'[a]}]", float(grad[vi, a]), ref))
    return worst


def test_vjp_vertices_two_ray_matches_fd(device):
    _say("case: vertex VJP two-ray 16x16 at 2 m, 3.5 GHz, los + ground reflection")
    scene = workloads.two_ray_scene(shape=(16, 16), spacing=2.0, extent=40.0)
    worst = _vertex_case(device, scene, {"los", "reflection"}, 1, 3.5e9)
    assert worst <= REL_TOL


def test_vjp_vertices_urban_canyon_matches_fd(device):
    _say("case: vertex VJP urban canyon seed 2, 16x16 at 20 m, depth 2, los+refl, 3.5 GHz")
    scene = workloads.urban_canyon_scene(seed=2, shape=(16, 16), spacing=20.0)
    worst = _vertex_case(device, scene, {"los", "reflection"}, 2, 3.5e9)
    assert worst <= REL_TOL


@pytest.mark.skip(reason="edge (diffraction) vertex gradients are switched off: their kernel stopped with a GPU "
                         "memory aperture violation of unidentified cause")
def test_vjp_vertices_with_diffraction_matches_fd(device):
    _say("case: vertex VJP urban canyon seed 2, 16x16 at 20 m, depth 2, los+refl+diff, 3.5 GHz")
    scene = workloads.urban_canyon_scene(seed=2, shape=(16, 16), spacing=20.0)
    worst = _vertex_case(device, scene, {"los", "reflection", "diffraction"}, 2, 3.5e9)
    assert worst <= REL_TOL


def test_vertex_gradient_with_diffraction_is_refused(device):
    _say("case: vertex VJP request with diffraction enabled raises (no kernel launch)")
    scene = workloads.urban_canyon_scene(seed=2, shape=(4, 4), spacing=20.0)
    run = run_gpu(scene, {"los", "reflection", "diffraction"}, 1, 3.5e9, device, topology_capacity=64)
    m = run.compiled.materials
    verts = torch.tensor(run.compiled.vertices, device=device, requires_grad=True)
    outs, _ = radiomap_vjp_outputs(run.dscene, run.inputs, eps_r=torch.tensor(m.eps_r, device=device),
                                   sigma_e=torch.tensor(m.sigma_e, device=device),
                                   frequency=torch.tensor(3.5e9, dtype=torch.float32, device=device),
                                   tx_position=torch.tensor(run.compiled.endpoints.tx_position, device=device),
                                   material_count=m.count, vertices=verts)
    with pytest.raises(RuntimeError, match="diffraction"):
        outs[0].sum().backward()


def test_vjp_vertices_transmission_walls_match_fd(device):
    _say("case: vertex VJP thin-sheet transmission through a tilted wall + ground, 3 GHz")
    from witwin.core import Mesh, PhysicalMaterial, Scene, Structure

    v = np.array([[3.0, -6, 0], [3.4, 6, 0], [3.2, -6, 8], [3.6, 6, 8]], np.float32)
    wall = Structure(Mesh(torch.from_numpy(v), torch.tensor([[0, 1, 2], [1, 3, 2]])),
                     PhysicalMaterial(eps_r=4.4, sigma_e=0.02, thickness_m=0.12), surface_id=1)
    ground = PhysicalMaterial(eps_r=5.31, sigma_e=0.0326, thickness_m=5.0)
    scene = Scene(structures=[workloads.ground_structure(-50, -50, 50, 50, ground), wall],
                  endpoints=[workloads.transmitter((0.0, 0.3, 2.0)), workloads.grid((8.0, -4.0, 1.5), (8, 8), 1.0)])
    worst = _vertex_case(device, scene, {"los", "reflection", "transmission"}, 1, 3.0e9, picks=3)
    assert worst <= REL_TOL


def test_mean_db_step_vertex_gradient_matches_autograd_and_replays(device):
    _say("case: MeanDbStep(vertices=True) on the two-ray 16x16 map: eager, graph replays, autograd")
    import math

    from witwin.channel.kernels.deterministic import MeanDbStep

    scene = workloads.two_ray_scene(shape=(16, 16), spacing=2.0, extent=40.0)
    run = run_gpu(scene, {"los", "reflection"}, 1, 3.5e9, device, topology_capacity=64)
    m = run.compiled.materials
    step = MeanDbStep(run.dscene, run.inputs, m.count, vertices=True)
    step.launch()
    torch.cuda.synchronize()
    eager = step.grad_vertices.clone()
    step.capture()
    for _ in range(3):
        step.step()
        torch.cuda.synchronize()
        assert torch.allclose(step.grad_vertices, eager, rtol=1e-6, atol=1e-12)
    verts = torch.tensor(run.compiled.vertices, device=device, requires_grad=True)
    outs, _ = radiomap_vjp_outputs(run.dscene, run.inputs, eps_r=torch.tensor(m.eps_r, device=device),
                                   sigma_e=torch.tensor(m.sigma_e, device=device),
                                   frequency=torch.tensor(3.5e9, dtype=torch.float32, device=device),
                                   tx_position=torch.tensor(run.compiled.endpoints.tx_position, device=device),
                                   material_count=m.count, vertices=verts)
    power = outs[0]
    lit = power > 0
    loss = (10.0 * torch.log10(power[lit])).sum() / max(int(lit.sum()), 1)
    loss.backward()
    torch.cuda.synchronize()
    scale = float(eager.abs().max())
    diff = float((verts.grad - eager).abs().max())
    _say(f"  max |autograd - step| {diff:.3e} (max |grad| {scale:.3e})")
    assert math.isfinite(scale) and scale > 0 and diff <= 1e-4 * scale
