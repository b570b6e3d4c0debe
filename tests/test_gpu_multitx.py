"""Several transmitters with their own scattered receiver sets in one launch
against one solve per transmitter: per-cell sums run in canonical order in
one lane, so every output of a cell is bit-identical whichever launch
evaluates it (the batched launch takes the per-transmitter-point-set paths of
the kernels, e.g. the queued edge visibility; the single solves the grid
paths)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from witwin.channel import workloads
from witwin.channel.components import component_bits
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def test_batched_transmitters_match_per_transmitter_solves(device):
    _say("case: 6 Tx with scattered street receivers (16..256 each), canyon, depth 2 + UTD + LoS, batched vs per Tx")
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(4, 4), spacing=2.0),
                             reference_frequency_hz=28e9)
    ds = device_scene(compiled, device)
    bits = component_bits({"los", "reflection", "diffraction"})
    tx_pos, sizes, rx_pos = workloads.batched_transmitters(seed=3, tx_count=6, max_rx=256)
    rx_pol = np.tile(np.array([[0, 0, 1]], np.float32), (rx_pos.shape[0], 1))
    tx_pol = np.tile(np.array([[0, 0, 1]], np.float32), (tx_pos.shape[0], 1))
    power = np.ones(tx_pos.shape[0], np.float32)
    batched = engine.solve_inputs(compiled, ds, components=bits, max_depth=2, coherent=True, frequency_hz=28e9,
                                  tx_pos=tx_pos, tx_pol=tx_pol, tx_power=power, rx_pos=rx_pos, rx_pol=rx_pol,
                                  rx_count=sizes)
    assert int(batched.view.rx_shared) == 0
    out = engine.radiomap_forward(ds, batched)
    torch.cuda.synchronize()
    names = ("power_total", "field_total", "component_power", "component_field", "diffraction_vector", "path_count")
    got = {k: getattr(out, k).cpu().numpy() for k in names}
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    total_paths = 0
    for t in range(tx_pos.shape[0]):
        sl = slice(int(offsets[t]), int(offsets[t + 1]))
        single = engine.solve_inputs(compiled, ds, components=bits, max_depth=2, coherent=True, frequency_hz=28e9,
                                     tx_pos=tx_pos[t:t + 1], tx_pol=tx_pol[t:t + 1], tx_power=power[t:t + 1],
                                     rx_pos=rx_pos[sl], rx_pol=rx_pol[sl])
        ref = engine.radiomap_forward(ds, single)
        torch.cuda.synchronize()
        for k in names:
            want = getattr(ref, k).cpu().numpy()
            have = got[k][..., sl] if k.startswith("component") else got[k][sl]
            np.testing.assert_array_equal(have, want, err_msg=f"tx {t} {k}")
        total_paths += int(ref.path_count.sum().item())
    _say(f"cells {int(sizes.sum())}, paths {total_paths}")
    assert total_paths > 0
