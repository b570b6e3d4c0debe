"""GPU tests of the coupled reflection-diffraction paths (wc_coupled_paths,
engine.coupled_fields, deterministic.solve(coupled_paths=True)) against the
FP64 oracle (oracle_coupled_paths, pinned on the CPU by image theory and
Fermat stationarity in test_cpu_coupled_oracle.py).

Scene: the reference's coupled wall-wedge geometry (tests/support/scenes.py:
218-264), a reflector at z = 0 and an L-shaped wedge, with receiver grids
around it; perfectly conducting and dielectric variants.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

import oracle
from witwin.channel import workloads
from witwin.channel.components import component_bits
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

pytestmark = pytest.mark.gpu
F = 3.0e9


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _grid_scene(material=None, shape=(14, 14)):
    return workloads.coupled_wall_wedge_scene(material=material, tx=(5.5, 0.3, 3.0), grid_shape=shape,
                                              grid_origin=(-3.0, -3.2, 2.6), grid_spacing=0.6)


def _run(device, scene):
    compiled = compile_scene(scene, reference_frequency_hz=F)
    dscene = device_scene(compiled, device)
    bits = component_bits({"los", "reflection", "diffraction"})
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=2, coherent=True, frequency_hz=F)
    co = engine.coupled_fields(dscene, inputs)
    torch.cuda.synchronize()
    return compiled, inputs, co


@pytest.mark.parametrize("material", ["pec", "concrete"])
def test_coupled_cells_match_oracle(device, material):
    from witwin.core import PhysicalMaterial

    mat = None if material == "pec" else PhysicalMaterial(eps_r=5.31, sigma_e=0.0326, thickness_m=0.2)
    _say(f"case: coupled R->D / D->R / D->D on the wall-wedge scene ({material}), 14x14 receivers, 3 GHz")
    compiled, inputs, co = _run(device, _grid_scene(mat))
    o = oracle.Oracle(compiled)
    ep = compiled.endpoints
    tx = ep.tx_position[0]
    field = co.field.cpu().numpy()
    count = co.count.cpu().numpy()
    count_dd = co.count_dd.cpu().numpy()
    scale = max(float(np.abs(field).max()), 1e-30)
    mismatched = 0
    worst = 0.0
    kinds = {3: 0, 4: 0, 7: 0}
    unstable = 0
    for c in range(inputs.total_cells):
        paths = o.coupled_paths(tx=tx, rx=ep.rx_position[c], frequency=F)
        for p in paths:
            kinds[p.component] += 1
        if len(paths) != int(count[c]) or sum(p.component == 7 for p in paths) != int(count_dd[c]):
            mismatched += 1
            # a float32 decision at its boundary (a reflection point on a face
            # edge, a grazing leg): the oracle's own path set changes when the
            # receiver moves by 0.2 mm along some axis
            sig = sorted((p.component, p.edge, p.prim[0]) for p in paths)
            moved = [sorted((p.component, p.edge, p.prim[0]) for p in
                            o.coupled_paths(tx=tx, rx=np.asarray(ep.rx_position[c], np.float64) + d, frequency=F))
                     for d in np.concatenate([np.eye(3), -np.eye(3)]) * 2e-4]
            boundary = any(m != sig for m in moved)
            unstable += 0 if boundary else 1
            _say(f"  cell {c} rx {np.round(ep.rx_position[c], 3)}: oracle {len(paths)} paths "
                 f"({sum(p.component == 7 for p in paths)} D->D) vs GPU {int(count[c])} ({int(count_dd[c])}); "
                 f"oracle set changes under 0.2 mm moves: {boundary}; oracle paths {sig}")
            continue
        ref = sum(complex(*p.path_field) for p in paths)
        err = abs(field[c] - ref)
        worst = max(worst, err / scale)
        assert err <= 2e-4 * abs(ref) + 1e-6 * scale, f"cell {c}: {field[c]} vs {ref}"
    _say(f"  paths R->D {kinds[3]}, D->R {kinds[4]}, D->D {kinds[7]}; cells with another path set: "
         f"{mismatched} of {inputs.total_cells}; worst field error / max field {worst:.2e}")
    assert kinds[3] > 0 and kinds[4] > 0 and kinds[7] > 0
    assert mismatched <= 0.05 * inputs.total_cells
    assert unstable == 0, "a topology-stable cell has another coupled path set on the GPU"


def test_solve_adds_the_coupled_slot(device):
    from witwin.channel.deterministic import Config, solve

    _say("case: deterministic.solve(coupled_paths=True): coupled slot joins the coherent and incoherent totals")
    scene = _grid_scene(shape=(10, 10))
    comps = {"los", "reflection", "diffraction"}
    for coherent in (True, False):
        base = solve(scene, Config(components=comps, max_depth=2, coherent=coherent, return_field=True),
                     reference_frequency_hz=F)
        res = solve(scene, Config(components=comps, max_depth=2, coherent=coherent, return_field=True,
                                  coupled_paths=True), reference_frequency_hz=F)
        torch.cuda.synchronize()
        assert "coupled" in res.component_power and "coupled" not in base.component_power
        cf = res.component_fields["coupled"].cpu().numpy().astype(np.complex128)
        cp = res.component_power["coupled"].cpu().numpy()
        if coherent:
            expect = np.abs(base.field.cpu().numpy().astype(np.complex128) + cf) ** 2
            np.testing.assert_allclose(cp, np.abs(cf) ** 2, rtol=1e-5, atol=1e-30)
        else:
            expect = base.path_gain.cpu().numpy().astype(np.float64) + cp
        np.testing.assert_allclose(res.path_gain.cpu().numpy(), expect, rtol=2e-5, atol=1e-30)
        meta = res.metadata["coupled_paths"]
        assert meta["requested"] and meta["double_diffraction"]["component_id"] == 7
        counts = res.metadata["counts"]["components"]
        assert counts["coupled"] > 0 and counts["coupled_double_diffraction"] <= counts["coupled"]
        assert float(np.abs(cf).max()) > 0.0


def test_candidate_budget_streams_blocks_and_raises_per_cell(device):
    _say("case: coupled candidate budget: small blocks give the same sums; a cell over the budget raises")
    scene = _grid_scene(shape=(6, 6))
    compiled = compile_scene(scene, reference_frequency_hz=F)
    dscene = device_scene(compiled, device)
    bits = component_bits({"los", "reflection", "diffraction"})
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=2, coherent=True, frequency_hz=F)
    per_cell = engine.coupled_candidates_per_cell(dscene)
    whole = engine.coupled_fields(dscene, inputs)
    blocks = engine.coupled_fields(dscene, inputs, candidate_limit=per_cell * 5)
    torch.cuda.synchronize()
    assert torch.equal(torch.view_as_real(whole.field), torch.view_as_real(blocks.field))
    assert torch.equal(whole.count, blocks.count)
    with pytest.raises(RuntimeError, match="coupled_candidate_limit"):
        engine.coupled_fields(dscene, inputs, candidate_limit=per_cell - 1)
