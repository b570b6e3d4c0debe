"""GPU parity beyond the sampled checks of test_gpu_parity: the full
256 x 256 urban-canyon map (4,096 sampled cells with the coherent phase),
reflection depths 4 and 5, the thin-sheet transmission fixtures of the
reference (vacuum wall, lossy wall, two walls, line-of-sight exclusivity,
penetration overflow), the grid diffraction override against the oracle's
vector sum, a forced capacity regrow, and the discovery chain set against
the oracle's own rays (all 262,144 of them)."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import DB_TOL, PHASE_TOL, db_error, phase_error, run_gpu, run_oracle
from witwin.channel import workloads
from witwin.channel.propagation import engine

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _cell_checks(run, idx, ref, label):
    """Path counts on topology-stable cells; power within 0.01 dB where the
    cell is not a deep null (|sum F|^2 >= 1e-3 sum |F|^2), coherent phase
    within 1e-3 rad where |sum F|^2 >= 1e-2 sum |F|^2. Returns the worst
    errors."""
    rec = ref["record"]
    stable = ~ref["ambiguous"].astype(bool)
    counts = run.out.path_count.cpu().numpy()[idx]
    p_gpu = run.out.power_total.cpu().numpy()[idx]
    f_gpu = run.out.field_total.cpu().numpy()[idx]
    bad = np.nonzero(stable & (counts != ref["counts"]))[0]
    assert bad.size == 0, f"{label}: path counts differ on {bad.size} stable cells"
    incoherent = np.array([sum(p.gain for p in cell) for cell in ref["paths"]])
    cond = rec[:, 0] / np.maximum(incoherent, 1e-300)
    err = db_error(p_gpu, rec[:, 0])
    well = stable & (cond > 1e-3) & (rec[:, 0] > 1e-25)
    phased = stable & (cond > 1e-2) & (rec[:, 0] > 1e-25)
    ph = phase_error(f_gpu, rec[:, 1] + 1j * rec[:, 2])
    worst_db = float(err[well].max()) if well.any() else 0.0
    worst_ph = float(ph[phased].max()) if phased.any() else 0.0
    _say(f"{label}: {len(idx)} cells, {int((~stable).sum())} ambiguous, {int(well.sum())} checked in dB "
         f"(worst {worst_db:.2e}), {int(phased.sum())} in phase (worst {worst_ph:.2e} rad)")
    assert worst_db <= DB_TOL
    assert worst_ph <= PHASE_TOL
    return worst_db, worst_ph


def test_canyon_full_map_4096_cells_with_phase(device):
    _say("case: urban canyon seed 2, 256x256 grid at 2 m, depth 3, los+refl+diff, 28 GHz; 4096 sampled cells")
    scene = workloads.urban_canyon_scene(seed=2, shape=(256, 256), spacing=2.0)
    run = run_gpu(scene, {"los", "reflection", "diffraction"}, 3, 28e9, device)
    rng = np.random.default_rng(11)
    idx = np.sort(rng.choice(run.inputs.total_cells, 4096, replace=False))
    ref = run_oracle(run, 0, idx, paths_cap=512, perturb=2e-4)
    assert int(ref["counts"].max()) <= 512
    assert (~ref["ambiguous"].astype(bool)).sum() >= 0.85 * len(idx)
    _cell_checks(run, idx, ref, "canyon 256x256")


@pytest.mark.parametrize("depth", [4, 5])
def test_reflection_depths_4_and_5_match_oracle(device, depth):
    _say(f"case: urban canyon seed 2, 32x32 grid at 10 m, reflection to depth {depth} + LoS, 28 GHz; 256 cells")
    scene = workloads.urban_canyon_scene(seed=2, shape=(32, 32), spacing=10.0)
    run = run_gpu(scene, {"los", "reflection"}, depth, 28e9, device)
    groups = run.inputs.chains.groups_host
    per_depth = [(groups >= 0).sum(axis=1) == d for d in range(1, depth + 1)]
    _say(f"chains per depth {[int(m.sum()) for m in per_depth]}")
    assert int(per_depth[-1].sum()) > 0
    rng = np.random.default_rng(depth)
    idx = np.sort(rng.choice(run.inputs.total_cells, 256, replace=False))
    ref = run_oracle(run, 0, idx, paths_cap=512, perturb=2e-4)
    assert int(ref["counts"].max()) <= 512
    deep = sum(1 for cell in ref["paths"] for p in cell if p.depth == depth)
    _say(f"oracle paths of depth {depth} on the sampled cells: {deep}")
    assert deep > 0
    _cell_checks(run, idx, ref, f"depth {depth}")


# ----------------------------------------------------------------- transmission
# reference tests/deterministic/test_transmission.py:35-260 (walls of
# tests/support/scenes.py:150-169: 8 x 8 m sheets in x = const, 3 GHz,
# transmitter at the origin)

def _wall(x, material, sid):
    from witwin.core import Mesh, Structure

    v = torch.tensor([[x, -4.0, -4.0], [x, 4.0, -4.0], [x, -4.0, 4.0], [x, 4.0, 4.0]])
    return Structure(Mesh(v, torch.tensor([[0, 1, 2], [1, 3, 2]])), material, name=f"wall-{sid}", surface_id=sid)


def _vacuum(thickness=0.3):
    from witwin.core import MaterialLayer, PhysicalMaterial

    return PhysicalMaterial(layers=(MaterialLayer(thickness_m=thickness, eps_r=1.0),), name="vacuum-wall")


def _lossy(name="lossy-wall"):
    from witwin.core import MaterialLayer, PhysicalMaterial

    return PhysicalMaterial(layers=(MaterialLayer(thickness_m=0.1, eps_r=4.0, sigma_e=0.05),), name=name)


def _solve(structures, rx, components, max_depth=1, coherent=True):
    from witwin.channel.deterministic import Config, solve
    from witwin.core import Scene

    scene = Scene(structures=structures, endpoints=[workloads.transmitter((0.0, 0.0, 0.0)), workloads.receiver(rx)])
    return solve(scene, Config(components=set(components), max_depth=max_depth, coherent=coherent),
                 reference_frequency_hz=3.0e9)


@pytest.mark.parametrize("rx", [(5.0, 0.0, 0.0), (5.0, 5.0, 0.0)])
def test_transmission_vacuum_wall_equals_line_of_sight(device, rx):
    _say(f"case: vacuum wall at x = 2.5, rx {rx}")
    wall = _solve([_wall(2.5, _vacuum(), 1)], rx, {"transmission"})
    empty = _solve([], rx, {"los"})
    ratio = wall.component_fields["transmission"] / empty.component_fields["los"]
    assert torch.abs(ratio - 1.0).max().item() <= 1.0e-4
    torch.testing.assert_close(wall.component_power["transmission"], empty.component_power["los"], rtol=1e-4,
                               atol=0.0)
    assert wall.metadata["transmission"] == {"thin_sheet_straight_path_approximation": True,
                                             "group_delay": "geometric"}


@pytest.mark.parametrize(("rx", "cos_theta"), [((5.0, 0.0, 0.0), 1.0), ((5.0, 5.0, 0.0), math.cos(math.pi / 4))])
def test_transmission_lossy_wall_matches_layer_stack(device, rx, cos_theta):
    _say(f"case: lossy wall (eps 4, sigma 0.05, 0.1 m) at x = 2.5, rx {rx}")
    wall = _solve([_wall(2.5, _lossy(), 1)], rx, {"transmission"})
    empty = _solve([], rx, {"los"})
    expected = abs(oracle.stack_t(cos_theta, [0.1], [4.0], [0.05], [1.0], 3.0e9, 0))
    assert expected < 0.999
    observed = (wall.component_fields["transmission"].abs() / empty.component_fields["los"].abs()).item()
    assert observed == pytest.approx(expected, rel=1e-4)
    power_ratio = (wall.component_power["transmission"] / empty.component_power["los"]).item()
    assert power_ratio == pytest.approx(expected ** 2, rel=2e-4)


def test_transmission_two_walls_are_the_product(device):
    _say("case: lossy wall at x = 2 and vacuum wall at x = 3, depth 2")
    rx = (5.0, 0.0, 0.0)
    a, b = _wall(2.0, _lossy("wall-a"), 1), _wall(3.0, _vacuum(), 2)
    los = _solve([], rx, {"los"}).component_fields["los"]
    ra = _solve([a], rx, {"transmission"}).component_fields["transmission"] / los
    rb = _solve([b], rx, {"transmission"}).component_fields["transmission"] / los
    both = _solve([a, b], rx, {"transmission"}, max_depth=2)
    rab = both.component_fields["transmission"] / los
    assert torch.abs(rab - ra * rb).max().item() <= 1.0e-4


def test_transmission_overflow_fails_loudly_and_excludes_los(device):
    _say("case: two vacuum walls with max_depth 1 (overflow), and line-of-sight exclusivity")
    rx = (5.0, 0.0, 0.0)
    with pytest.raises(RuntimeError, match="penetration overflow"):
        _solve([_wall(2.0, _vacuum(), 1), _wall(3.0, _vacuum(), 2)], rx, {"transmission"})
    blocked = _solve([_wall(2.5, _lossy(), 1)], rx, {"los", "transmission"}, coherent=False)
    assert torch.count_nonzero(blocked.component_power["los"]) == 0
    assert bool((blocked.component_power["transmission"] > 0).all())
    torch.testing.assert_close(blocked.path_gain, blocked.component_power["transmission"])


def test_transmission_radiomap_matches_oracle(device):
    _say("case: layered walls (1-3 layers) across a street grid, transmission + LoS depth 3, 28 GHz")
    from witwin.core import MaterialLayer, PhysicalMaterial, Scene

    stacks = [
        PhysicalMaterial(layers=(MaterialLayer(thickness_m=0.12, eps_r=5.31, sigma_e=0.0326),), name="one"),
        PhysicalMaterial(layers=(MaterialLayer(thickness_m=0.01, eps_r=6.27, sigma_e=0.0043),
                                 MaterialLayer(thickness_m=0.05, eps_r=1.0),
                                 MaterialLayer(thickness_m=0.01, eps_r=6.27, sigma_e=0.0043)), name="three"),
        PhysicalMaterial(layers=(MaterialLayer(thickness_m=0.02, eps_r=2.94, sigma_e=0.0116),
                                 MaterialLayer(thickness_m=0.1, eps_r=1.99, sigma_e=0.0047)), name="two"),
    ]
    walls = [_wall(x, m, k + 1) for k, (x, m) in enumerate(zip((2.0, 3.5, 5.0), stacks))]
    scene = Scene(structures=walls, endpoints=[workloads.transmitter((0.0, 0.3, 0.2)),
                                                workloads.grid((6.0, -3.0, -2.0), (16, 16), 0.4)])
    run = run_gpu(scene, {"los", "transmission"}, 3, 28e9, device)
    idx = np.arange(run.inputs.total_cells)
    ref = run_oracle(run, 0, idx, paths_cap=8, perturb=2e-4)
    slot = run.out.component_power[3].cpu().numpy()
    assert (slot > 0).sum() > 100
    _cell_checks(run, idx, ref, "layered walls")
    field = run.out.component_field[3].cpu().numpy()
    rec = ref["record"]
    lit = rec[:, 6] > 0
    assert phase_error(field[lit], rec[lit, 10] + 1j * rec[lit, 14]).max() <= PHASE_TOL


# ----------------------------------------------------------------- accumulation, capacity, discovery

def test_grid_diffraction_override_matches_oracle_vector_sum(device):
    _say("case: deterministic.solve grid diffraction power vs the oracle's per-axis vector sum")
    from witwin.channel.deterministic import Config, solve
    from witwin.channel.scene.compiler import compile as compile_scene

    scene = workloads.urban_canyon_scene(seed=2, shape=(24, 24), spacing=15.0)
    result = solve(scene, Config(components={"los", "reflection", "diffraction"}, max_depth=2),
                   reference_frequency_hz=28e9)
    gpu = result.component_power["diffraction"][0].transpose(0, 1).reshape(-1).cpu().numpy()  # [cols, rows] -> flat
    run = run_gpu(scene, {"los", "reflection", "diffraction"}, 2, 28e9, device)
    idx = np.arange(run.inputs.total_cells)
    ref = run_oracle(run, 0, idx, perturb=2e-4)
    dv = ref["record"][:, 15:21]
    vec = dv[:, 0] ** 2 + dv[:, 1] ** 2 + dv[:, 2] ** 2 + dv[:, 3] ** 2 + dv[:, 4] ** 2 + dv[:, 5] ** 2
    stable = ~ref["ambiguous"].astype(bool) & (vec > 1e-25)
    assert stable.sum() > 50
    err = db_error(gpu[stable], vec[stable])
    _say(f"diffraction override: {int(stable.sum())} cells, worst {err.max():.2e} dB")
    assert err.max() <= DB_TOL
    del compile_scene


def test_forced_capacity_regrow_reproduces_outputs(device):
    _say("case: canyon 32x32, topology capacity 1 forces the regrow")
    scene = workloads.urban_canyon_scene(seed=2, shape=(32, 32), spacing=8.0)
    run = run_gpu(scene, {"los", "reflection", "diffraction"}, 3, 28e9, device)
    small = engine.forward_with_topology(run.dscene, run.inputs, capacity=1)
    big = engine.forward_with_topology(run.dscene, run.inputs, capacity=4096)
    torch.cuda.synchronize()
    assert small.topology_capacity > 1 and int(small.failure.item()) == 0
    for k in ("power_total", "field_total", "path_count", "component_power"):
        assert torch.equal(getattr(small, k), getattr(big, k)), k
    counts = big.path_count.unsqueeze(1)
    cols = torch.arange(small.topology_capacity, device=device).unsqueeze(0)
    assert torch.equal(small.topology[cols < counts], big.topology[:, :small.topology_capacity][cols < counts])


def test_discovery_chain_set_equals_oracle_over_all_rays(device):
    """All 262,144 discovery rays of the canyon transmitter against the
    oracle's FP64 rays. Rays whose closest hit is decided by float32 rounding
    at a shared edge may pick a neighbouring group, so the chain sets may
    differ by the chains of those rays only; the radiomap built from the
    oracle's chain set must then have the same path count in every cell of
    the 256 x 256 map (an extra candidate is validated exactly per cell, a
    missing one would show as a missing path)."""
    _say("case: urban canyon, all 262,144 discovery rays to depth 3: GPU vs oracle chain sets")
    scene = workloads.urban_canyon_scene(seed=2, shape=(256, 256), spacing=2.0)
    run = run_gpu(scene, {"los", "reflection"}, 3, 28e9, device)
    tx = torch.as_tensor(run.compiled.endpoints.tx_position[:1]).to(device)
    gpu = engine.discover(run.dscene, tx, 3, 262_144)[0]
    ref = oracle.Oracle(run.compiled).discover_all(run.compiled.endpoints.tx_position[0], 262_144, 3)
    differ = ~(gpu == ref).all(axis=1)
    _say(f"rays identical: {int((~differ).sum())} / {len(differ)}")
    assert differ.sum() <= 3
    for d in (1, 2, 3):
        a = {tuple(r) for r in engine.unique_prefixes(gpu, d)}
        b = {tuple(r) for r in engine.unique_prefixes(ref, d)}
        explained = {tuple(r) for r in engine.unique_prefixes(gpu[differ], d)} | \
            {tuple(r) for r in engine.unique_prefixes(ref[differ], d)}
        _say(f"depth {d}: {len(a)} chains (GPU), {len(b)} (oracle), {len(a ^ b)} differ")
        assert (a ^ b) <= explained
    plan = engine.plan_chains(run.compiled, run.dscene, tx, 3, reflection=True, discovered=ref[None])
    alt = engine.solve_inputs(run.compiled, run.dscene, components=run.bits, max_depth=3, coherent=True,
                              frequency_hz=28e9, chains=plan)
    out = engine.radiomap_forward(run.dscene, alt)
    torch.cuda.synchronize()
    assert torch.equal(out.path_count, run.out.path_count)
    _say(f"path counts identical over {run.inputs.total_cells} cells with either chain set "
         f"({int(run.out.path_count.sum())} paths)")
