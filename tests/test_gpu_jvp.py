"""GPU tests of the forward-mode radiomap tangents (wc_radiomap_jvp,
RadiomapFunction.jvp, deterministic.solve(ad_mode="jvp")).

* Along a random direction of (first-layer eps_r and sigma of every used
  material, frequency, transmitter position) the tangent of
  L = sum_c w_c P_c must match the central difference of the FP64 oracle
  (the reference's AD check, tests/ad/_fd.py:20-62): urban canyon with UTD
  diffraction, and two layered walls with transmission.
* Forward and reverse mode agree: <g, J v> = <J^T g, v> for random g and v
  (the VJP kernel is pinned against the oracle separately).
* solve(ad_mode="jvp") under torch.autograd.forward_ad carries the same
  tangents to Result.path_gain.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

import oracle
from gpu_helpers import run_gpu, tx_chains
from test_gpu_vjp import _oracle_power
from witwin.channel import workloads
from witwin.channel.kernels.deterministic import radiomap_jvp, radiomap_vjp_outputs

pytestmark = pytest.mark.gpu
REL_TOL = 1.0e-2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _walls_scene():
    from witwin.core import MaterialLayer, Mesh, PhysicalMaterial, Scene, Structure

    def wall(x, mat, sid):
        v = np.array([[x, -6, 0], [x, 6, 0], [x, -6, 8], [x, 6, 8]], np.float32)
        return Structure(Mesh(torch.from_numpy(v), torch.tensor([[0, 1, 2], [1, 3, 2]])), mat, surface_id=sid)

    brick = PhysicalMaterial(layers=(MaterialLayer(0.12, 4.4, 0.02), MaterialLayer(0.05, 2.2, 0.001)))
    glass = PhysicalMaterial(eps_r=6.27, sigma_e=0.0043, thickness_m=0.02)
    ground = PhysicalMaterial(eps_r=5.31, sigma_e=0.0326, thickness_m=5.0)
    return Scene(structures=[workloads.ground_structure(-50, -50, 50, 50, ground), wall(3.0, brick, 1),
                             wall(6.5, glass, 2)],
                 endpoints=[workloads.transmitter((0.0, 0.3, 2.0)), workloads.grid((8.0, -4.0, 1.5), (8, 8), 1.0)])


def _setup(device, scene, components, depth, frequency):
    run = run_gpu(scene, components, depth, frequency, device, topology_capacity=64)
    compiled = run.compiled
    m = compiled.materials
    eps = torch.tensor(m.eps_r, device=device)
    sig = torch.tensor(m.sigma_e, device=device)
    freq = torch.tensor(frequency, dtype=torch.float32, device=device)
    txp = torch.tensor(compiled.endpoints.tx_position, device=device)
    outs, state = radiomap_vjp_outputs(run.dscene, run.inputs, eps_r=eps, sigma_e=sig, frequency=freq,
                                       tx_position=txp, material_count=m.count)
    torch.cuda.synchronize()
    assert not bool((state.out.failure != 0).any()), "topology capacity exceeded"
    return run, state


def _direction(compiled, frequency, seed):
    """One step per parameter kind, sized like the VJP test's difference steps."""
    rng = np.random.default_rng(seed)
    m = compiled.materials
    used = np.array([(compiled.face_material == i).any() for i in range(m.count)])
    d_eps = rng.uniform(-1, 1, m.count) * 1e-3 * used
    d_sig = rng.uniform(-1, 1, m.count) * np.maximum(m.sigma_e, 1e-3) * 1e-3 * used
    d_f = float(rng.uniform(-1, 1)) * frequency * 1e-7
    d_tx = rng.uniform(-1, 1, 3) * 1e-4
    return d_eps, d_sig, d_f, d_tx


def _jvp_case(device, scene, components, depth, frequency):
    run, state = _setup(device, scene, components, depth, frequency)
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
    w = w / max(float((w * base["record"][:, 0]).sum()), 1e-300)
    m = compiled.materials
    worst = 0.0
    for seed in (1, 2):
        d_eps, d_sig, d_f, d_tx = _direction(compiled, frequency, seed)
        t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=device)  # noqa: E731
        d_power = radiomap_jvp(state, t(d_eps), t(d_sig), torch.tensor(d_f, device=device), t(d_tx))[0]
        torch.cuda.synchronize()
        got = float((d_power.double().cpu().numpy() * w).sum())

        def L(sign):
            eps = m.eps_r.astype(np.float64) + sign * d_eps
            sig = m.sigma_e.astype(np.float64) + sign * d_sig
            tx = ep.tx_position[0].astype(np.float64) + sign * d_tx
            return float((w * _oracle_power(compiled, run, idx, eps=eps, sigma=sig, frequency=frequency + sign * d_f,
                                            tx=tx)).sum())

        fd = (L(1.0) - L(-1.0)) / 2.0
        rel = abs(got - fd) / max(abs(fd), 1e-6)
        _say(f"  direction {seed}: jvp {got:.6e} fd {fd:.6e} rel {rel:.2e}")
        worst = max(worst, rel)
    return worst


def test_jvp_urban_canyon_matches_fd(device):
    _say("case: JVP urban canyon seed 2, 24x24 at 15 m, depth 2, los+refl+diff, 28 GHz")
    scene = workloads.urban_canyon_scene(seed=2, shape=(24, 24), spacing=15.0)
    assert _jvp_case(device, scene, {"los", "reflection", "diffraction"}, 2, 28e9) <= REL_TOL


def test_jvp_transmission_walls_match_fd(device):
    _say("case: JVP thin-sheet transmission through two layered walls + ground, 3 GHz")
    assert _jvp_case(device, _walls_scene(), {"los", "reflection", "transmission"}, 2, 3.0e9) <= REL_TOL


@pytest.mark.parametrize("which", ["canyon", "walls"])
def test_jvp_and_vjp_are_adjoint(device, which):
    """<g, J v> = <J^T g, v> for the power, component-power and diffraction
    power outputs with random weights g and a random direction v."""
    if which == "canyon":
        _say("case: JVP/VJP adjointness, urban canyon 32x32 at 10 m, depth 3 + UTD, 28 GHz")
        scene = workloads.urban_canyon_scene(seed=2, shape=(32, 32), spacing=10.0)
        run, state = _setup(device, scene, {"los", "reflection", "diffraction"}, 3, 28e9)
        frequency = 28e9
    else:
        _say("case: JVP/VJP adjointness, layered walls, depth 2 + transmission, 3 GHz")
        run, state = _setup(device, _walls_scene(), {"los", "reflection", "transmission"}, 2, 3.0e9)
        frequency = 3.0e9
    compiled = run.compiled
    m = compiled.materials
    cells = run.inputs.total_cells
    rng = np.random.default_rng(5)
    out = state.out
    scale_p = 1.0 / max(float(out.power_total.abs().max()), 1e-30)
    g_p = torch.as_tensor(rng.uniform(-1, 1, cells) * scale_p, dtype=torch.float32, device=device)
    g_c = torch.as_tensor(rng.uniform(-1, 1, (4, cells)) * scale_p, dtype=torch.float32, device=device)
    g_d = torch.as_tensor(rng.uniform(-1, 1, cells) * scale_p, dtype=torch.float32, device=device)
    eps = torch.tensor(m.eps_r, device=device, requires_grad=True)
    sig = torch.tensor(m.sigma_e, device=device, requires_grad=True)
    freq = torch.tensor(frequency, dtype=torch.float32, device=device, requires_grad=True)
    txp = torch.tensor(compiled.endpoints.tx_position, device=device, requires_grad=True)
    outs, st = radiomap_vjp_outputs(run.dscene, run.inputs, eps_r=eps, sigma_e=sig, frequency=freq, tx_position=txp,
                                    material_count=m.count)
    loss = (outs[0] * g_p).sum() + (outs[2] * g_c).sum() + (outs[4] * g_d).sum()
    loss.backward()
    d_eps, d_sig, d_f, d_tx = _direction(compiled, frequency, 9)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=device)  # noqa: E731
    dp, _, dcp, _, ddp = radiomap_jvp(st, t(d_eps), t(d_sig), torch.tensor(d_f, device=device), t(d_tx))
    torch.cuda.synchronize()
    forward = float((dp.double() * g_p.double()).sum() + (dcp.double() * g_c.double()).sum()
                    + (ddp.double() * g_d.double()).sum())
    reverse = float((eps.grad.double().cpu() * torch.as_tensor(d_eps)).sum()
                    + (sig.grad.double().cpu() * torch.as_tensor(d_sig)).sum()
                    + float(freq.grad) * d_f + (txp.grad.double().cpu()[0] * torch.as_tensor(d_tx)).sum())
    rel = abs(forward - reverse) / max(abs(reverse), 1e-12)
    _say(f"  <g, Jv> {forward:.6e}  <J^T g, v> {reverse:.6e}  rel {rel:.2e}")
    assert rel <= 2e-3


def test_solve_ad_mode_jvp_carries_tangents(device):
    _say("case: deterministic.solve(ad_mode='jvp') under forward AD, ground reflection, eps_r tangent")
    from witwin.channel.deterministic import Config, solve
    from witwin.core import PhysicalMaterial, Scene

    scene_of = lambda eps: Scene(  # noqa: E731
        structures=[workloads.ground_structure(-1e4, -1e4, 1e4, 1e4,
                                               PhysicalMaterial(eps_r=eps, sigma_e=0.0326, thickness_m=5.0))],
        endpoints=[workloads.transmitter((0.0, 0.0, 10.0)), workloads.grid((-15.5, -15.5, 1.5), (16, 16), 2.0)])
    cfg = Config(components={"los", "reflection"}, max_depth=1)
    with fwAD.dual_level():
        eps = fwAD.make_dual(torch.tensor(5.31), torch.tensor(1.0))
        res = solve(scene_of(eps), Config(components=cfg.components, max_depth=1, ad_mode="jvp"),
                    reference_frequency_hz=3.5e9)
        primal, tangent = fwAD.unpack_dual(res.path_gain)
        assert tangent is not None, "path_gain carries no forward-mode tangent"
        tangent = tangent.clone()
    torch.cuda.synchronize()
    # the same derivative by reverse mode
    eps_r = torch.tensor(5.31, requires_grad=True)
    res_v = solve(scene_of(eps_r), Config(components=cfg.components, max_depth=1, ad_mode="vjp"),
                  reference_frequency_hz=3.5e9)
    w = torch.rand(res_v.path_gain.shape, generator=torch.Generator().manual_seed(3)).to(res_v.path_gain.device)
    (res_v.path_gain * w).sum().backward()
    fwd = float((tangent * w).sum())
    rev = float(eps_r.grad)
    _say(f"  forward {fwd:.6e} reverse {rev:.6e}")
    assert torch.isfinite(tangent).all()
    assert abs(fwd - rev) <= 1e-3 * max(abs(rev), 1e-20)
