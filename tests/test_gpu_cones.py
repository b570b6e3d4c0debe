"""GPU tests of the chains' receiver cones (wc_chain_cones + the cone test in
the chain culls): they may only drop (tile, chain) pairs that the exact
validity test rejects for every receiver of the tile, so every radiomap
output must be bit-identical with and without them: the urban canyon to
depth 4, the heightfield to depth 2, and four transmitters of the scaled city
to depth 5 through the batched solve (supertile cull + segment kernels)."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from witwin.channel import workloads
from witwin.channel.components import component_bits
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _outputs(out):
    return (out.power_total, out.field_total, out.component_power, out.component_field, out.path_count)


def _same(a, b):
    for x, y in zip(_outputs(a), _outputs(b)):
        assert torch.equal(x, y)
    cap = a.topology.shape[1]
    cols = torch.arange(cap, device=a.topology.device)[None, :]
    valid = cols < a.path_count[:, None]
    assert torch.equal(a.topology[valid], b.topology[valid])


@pytest.mark.parametrize("which", ["canyon", "heightfield"])
def test_chain_cones_are_bitwise_neutral(device, which):
    if which == "canyon":
        _say("case: urban canyon seed 2, 64x64 grid at 6 m, depth 4 + LoS + UTD, 28 GHz; cones on vs off")
        scene = workloads.urban_canyon_scene(seed=2, shape=(64, 64), spacing=6.0)
        comps, depth, freq = {"los", "reflection", "diffraction"}, 4, 28e9
    else:
        _say("case: heightfield seed 4, 64x64 grid, depth 2 + LoS, 3.5 GHz; cones on vs off")
        scene = workloads.heightfield_scene(seed=4, size=64, boxes=16, shape=(64, 64))
        comps, depth, freq = {"los", "reflection"}, 2, 3.5e9
    compiled = compile_scene(scene, reference_frequency_hz=freq)
    dscene = device_scene(compiled, device)
    bits = component_bits(comps)
    base = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True, frequency_hz=freq)
    runs = []
    for cones in (False, True):
        inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                     frequency_hz=freq, chains=base.chains, cones=cones)
        assert bool(inputs.view.chain_cones) == cones
        runs.append(engine.forward_with_topology(dscene, inputs, capacity=512))
    torch.cuda.synchronize()
    cones_t = base.chains.cones
    planes = cones_t[:, 3].cpu().numpy()
    _say(f"{base.chains.count} chains, {int((planes > 0).sum())} bounded, {int((planes == 1).sum())} empty; "
         f"{int(runs[0].path_count.sum())} paths")
    _same(runs[0], runs[1])


def test_chain_cones_neutral_on_scaled_city_batch(device):
    _say("case: scaled city seed 5, transmitters 0-3 batched, depth 5 + UTD + transmission, 28 GHz; cones on vs off")
    from witwin.channel.propagation import batching

    city = workloads.scaled_city(seed=5)
    compiled = compile_scene(city.scene, reference_frequency_hz=city.frequencies[-1])
    dscene = device_scene(compiled, device)
    bits = component_bits({"los", "reflection", "diffraction", "transmission"})
    tx = np.asarray(city.tx[:4], np.float32)
    cnt = np.asarray(city.rx_count[:4])
    rx = np.asarray(city.rx[:int(cnt.sum())], np.float32)
    tx_t = torch.as_tensor(tx).to(device)
    seqs = batching.chain_sequences_batched(compiled, dscene, tx_t, 5, True)
    plan = engine.plan_from_device(dscene, tx_t, seqs)
    pol = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (rx.shape[0], 1))
    runs = []
    for cones in (False, True):
        inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=5, coherent=True,
                                     frequency_hz=city.frequencies[-1], tx_pos=tx, tx_pol=pol[:4],
                                     tx_power=np.ones(4, np.float32), rx_pos=rx, rx_pol=pol, rx_count=cnt,
                                     chains=plan, visibility_mode=1, cones=cones)
        runs.append(engine.forward_with_topology(dscene, inputs, capacity=256))
    torch.cuda.synchronize()
    planes = plan.cones[:, 3].cpu().numpy()
    _say(f"{plan.count} chains, {int((planes > 0).sum())} bounded, {int((planes == 1).sum())} empty; "
         f"{int(runs[0].path_count.sum())} paths over {int(cnt.sum())} cells")
    _same(runs[0], runs[1])
