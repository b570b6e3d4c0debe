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
