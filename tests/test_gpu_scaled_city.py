"""GPU tests of the scaled city (seed 5: 400 boxes on a 256 x 256 fBm
heightfield, 134,850 triangles, 8 materials; 5-bounce reflection + UTD +
transmission + line of sight at 3.5 and 28 GHz): a transmitter batch against
per-transmitter solves (bit-identical maps), and sampled cells of two
transmitters' grids against the FP64 oracle (brute-force visibility over
every triangle) at 0.01 dB and 1e-3 rad."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import DB_TOL, PHASE_TOL, db_error, phase_error
from witwin.channel import workloads
from witwin.channel.components import component_bits
from witwin.channel.propagation import batching
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

pytestmark = pytest.mark.gpu
TX = (0, 6)  # a 382 x 382 grid and a 66 x 66 grid


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def city_run(device):
    _say("case: scaled city seed 5, Tx 0 and 6, depth 5 refl + UTD + transmission + LoS, 3.5 and 28 GHz")
    city = workloads.scaled_city(seed=5)
    compiled = compile_scene(city.scene, reference_frequency_hz=28e9)
    dscene = device_scene(compiled, device)
    off = np.concatenate([[0], np.cumsum(city.rx_count)])
    tx = city.tx[list(TX)]
    cnt = city.rx_count[list(TX)]
    rx = np.concatenate([city.rx[int(off[i]):int(off[i + 1])] for i in TX], axis=0)
    bits = component_bits({"los", "reflection", "diffraction", "transmission"})
    seqs = batching.chain_sequences_batched(compiled, dscene, torch.as_tensor(tx).to(device), 5, True)
    run = batching.BatchedMeanDb(compiled, dscene, tx, rx, cnt, components=bits, max_depth=5,
                                 frequencies=city.frequencies, shape=batching.LaunchShape(1, 2), sequences=seqs)
    run.launch()
    torch.cuda.synchronize()
    return city, compiled, dscene, tx, rx, cnt, bits, seqs, run


def test_scaled_city_batch_matches_per_transmitter_solves(city_run):
    city, compiled, dscene, tx, rx, cnt, bits, seqs, run = city_run
    single = batching.BatchedMeanDb(compiled, dscene, tx, rx, cnt, components=bits, max_depth=5,
                                    frequencies=city.frequencies, shape=batching.LaunchShape(2, 1), sequences=seqs)
    single.reuse_visibility = False
    single.launch()
    torch.cuda.synchronize()
    for fi in range(len(city.frequencies)):
        for name in ("power_total", "field_total", "path_count"):
            for a, b in zip(run.maps(fi, name), single.maps(fi, name)):
                assert torch.equal(a, b), (fi, name)
    torch.testing.assert_close(run.grad_tx, single.grad_tx, rtol=1e-4, atol=1e-9)
    _say(f"paths {run.paths}; batch (tile loops, visibility reused across carriers) == per-Tx queued solves")


@pytest.mark.parametrize("fi", [0, 1])
def test_scaled_city_cells_match_oracle(city_run, fi):
    city, compiled, dscene, tx, rx, cnt, bits, seqs, run = city_run
    freq = city.frequencies[fi]
    orc = oracle.Oracle(compiled)
    off = np.concatenate([[0], np.cumsum(cnt)])
    rng = np.random.default_rng(50 + fi)
    checked = 0
    worst_db = worst_phase = 0.0
    for k in range(len(TX)):
        n = int(cnt[k])
        sample = np.sort(rng.choice(n, 32, replace=False))
        groups = np.concatenate([np.pad(x, ((0, 0), (0, 5 - x.shape[1])), constant_values=-1)
                                 for x in seqs[k] if x.size], axis=0).astype(np.int32)
        depth = (groups >= 0).sum(axis=1).astype(np.int32)
        ref = orc.radiomap(tx=tx[k], rx=rx[int(off[k]):int(off[k + 1])][sample], frequency=freq, components=bits,
                           max_depth=5, chains=groups, chain_depth=depth, paths_cap=256, perturb=2e-4)
        p_gpu = run.maps(fi, "power_total")[k].cpu().numpy()[sample]
        f_gpu = run.maps(fi, "field_total")[k].cpu().numpy()[sample]
        n_gpu = run.maps(fi, "path_count")[k].cpu().numpy()[sample]
        rec = ref["record"]
        stable = ~ref["ambiguous"].astype(bool)
        incoherent = np.array([sum(p.gain for p in cell) for cell in ref["paths"]])
        for j in np.nonzero(stable)[0]:
            assert n_gpu[j] == int(rec[j, 21]), (k, j, n_gpu[j], rec[j, 21])
            if rec[j, 0] > 0:
                err = float(db_error(p_gpu[j:j + 1], rec[j:j + 1, 0])[0])
                worst_db = max(worst_db, err)
                if rec[j, 0] >= 1e-2 * incoherent[j]:  # not a deep interference null
                    ph = float(phase_error(f_gpu[j:j + 1], rec[j:j + 1, 1] + 1j * rec[j:j + 1, 2])[0])
                    worst_phase = max(worst_phase, ph)
            checked += 1
        _say(f"tx {TX[k]} at {freq / 1e9:g} GHz: {int(stable.sum())} / {len(sample)} stable cells, "
             f"{int((rec[:, 0] > 0).sum())} lit")
    _say(f"checked {checked} cells: worst {worst_db:.2e} dB, worst phase {worst_phase:.2e} rad")
    assert checked >= 48
    assert worst_db <= DB_TOL
    assert worst_phase <= PHASE_TOL
