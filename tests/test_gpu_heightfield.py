"""GPU tests of the heightfield workload (seeded fBm terrain, 128 x 128
vertices at 2 m pitch, 40 boxes, 4 materials; 32,738 triangles): the
radiomap on sampled cells of the 256 x 256 grid against the FP64 oracle
(brute-force visibility over every triangle), and the fixed-topology
gradient with respect to the mesh vertices against oracle central
differences on a coarse grid over the same terrain."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from gpu_helpers import db_error, run_gpu, run_oracle
from test_gpu_vjp import REL_TOL, _vertex_case
from witwin.channel import workloads

pytestmark = pytest.mark.gpu
DB_TOL = 0.01


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_heightfield_radiomap_matches_oracle(device):
    _say("case: heightfield seed 4, 256x256 grid, depth 2, los+refl, 3.5 GHz; 64 sampled cells")
    scene = workloads.heightfield_scene(seed=4)
    run = run_gpu(scene, {"los", "reflection"}, 2, 3.5e9, device)
    chains = run.inputs.chains.count
    cells = run.inputs.total_cells
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(cells, 64, replace=False))
    ref = run_oracle(run, 0, idx, paths_cap=256, perturb=2e-4)
    stable = ~ref["ambiguous"].astype(bool)
    counts = run.out.path_count.cpu().numpy()[idx]
    _say(f"chains {chains}; ambiguous cells {int((~stable).sum())} / {len(idx)}; "
         f"paths per sampled cell {counts.mean():.1f}")
    assert np.array_equal(counts[stable], ref["counts"][stable])
    p_gpu = run.out.power_total.cpu().numpy()[idx]
    rec = ref["record"]
    incoherent = np.array([sum(p.gain for p in cell) for cell in ref["paths"]])
    well = stable & (rec[:, 0] / np.maximum(incoherent, 1e-300) > 1e-3) & (rec[:, 0] > 1e-25)
    err = db_error(p_gpu, rec[:, 0])
    _say(f"heightfield: max dB error ({int(well.sum())} well-conditioned cells) {err[well].max():.3e}")
    assert err[well].max() <= DB_TOL


def test_heightfield_vertex_vjp_matches_fd(device):
    _say("case: vertex VJP heightfield seed 4, 12x12 grid, depth 1, los+refl, 3.5 GHz")
    scene = workloads.heightfield_scene(seed=4, shape=(12, 12))
    worst = _vertex_case(device, scene, {"los", "reflection"}, 1, 3.5e9)
    assert worst <= REL_TOL


def test_heightfield_tiled_launch_matches_plain_launch(device):
    """Many chains and one transmitter take the supertile chain cull (supertile
    receiver boxes, per-supertile chain masks, zero-filled chain masks) before
    the tile pass; the tiled launch must equal the plain launch, which traces
    every chain of every cell, bit for bit."""
    import torch

    from witwin.channel.components import component_bits
    from witwin.channel.propagation import engine
    from witwin.channel.scene.compiler import compile as compile_scene, device_scene

    _say("case: heightfield 64 x 64 grid (4 supertiles), LoS + depth-2 reflection, tiled vs plain launch")
    compiled = compile_scene(workloads.heightfield_scene(seed=4, shape=(64, 64)), reference_frequency_hz=3.5e9)
    ds = device_scene(compiled, device)
    inputs = engine.solve_inputs(compiled, ds, components=component_bits({"los", "reflection"}), max_depth=2,
                                 coherent=True, frequency_hz=3.5e9)
    words = (int(inputs.view.max_tx_chains) + 31) // 32
    _say(f"chains {inputs.chains.count}, mask words {words}, cells {inputs.total_cells}")
    assert words >= 64 and inputs.total_cells > 1024
    tiled = engine.radiomap_forward(ds, inputs, tiled=True)
    plain = engine.radiomap_forward(ds, inputs, tiled=False)
    torch.cuda.synchronize()
    for k in ("power_total", "field_total", "component_power", "component_field", "path_count"):
        a, b = getattr(tiled, k).cpu().numpy(), getattr(plain, k).cpu().numpy()
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
    assert int(tiled.path_count.sum().item()) > 0
