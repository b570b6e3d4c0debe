"""GPU tests of the batched explicit-path solve (urban canyon + three layered
street walls; 64 transmitters over the streets, each with its own street-point
receiver set of log-uniform size 16..4096; LoS, reflection and transmission to
depth 3 and first-order UTD diffraction at 28 GHz): canonical rows and pair
segmentation against the FP64 oracle on sampled pairs (bit-exact topology,
0.01 dB, 1e-3 rad), angles from the rows' directions, and the channel
frequency response on 1,024 subcarriers against float64."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import DB_TOL, PHASE_TOL, db_error, phase_error
from witwin.channel import workloads
from witwin.channel.components import component_bits
from witwin.channel.path import Config, solve_batched
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

pytestmark = pytest.mark.gpu
COMPONENTS = {"los", "reflection", "diffraction", "transmission"}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch(device):
    _say("case: batched path solve, canyon + 3 layered walls, 64 Tx with 16..4096 street points, depth 3, 28 GHz")
    w = workloads.batched_path_workload(seed=3, tx_count=64)
    compiled = compile_scene(w.scene, reference_frequency_hz=w.frequency)
    dscene = device_scene(compiled, device)
    bits = component_bits(COMPONENTS)
    t = w.tx.shape[0]
    pol = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (w.rx.shape[0], 1))
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=3, coherent=True,
                                 frequency_hz=w.frequency, tx_pos=w.tx, tx_pol=pol[:t], tx_power=np.ones(t, np.float32),
                                 rx_pos=w.rx, rx_pol=pol, rx_count=w.rx_count)
    paths = solve_batched(compiled, Config(components=COMPONENTS, max_depth=3), w.tx, w.rx, w.rx_count,
                          dscene=dscene, inputs=inputs)
    torch.cuda.synchronize()
    _say(f"pairs {paths.pair_count}, paths {paths.path_count}")
    return w, compiled, dscene, inputs, paths


def test_batched_paths_segmentation(batch):
    w, compiled, dscene, inputs, paths = batch
    t = w.tx.shape[0]
    assert paths.pair_count == int(w.rx_count.sum())
    tx_id = paths.rows["tx_id"].long()
    rx_id = paths.rows["rx_id"].long()
    key = rx_id * t + tx_id
    assert bool((key[1:] >= key[:-1]).all()), "rows are not in canonical pair order"
    pair_of_row = torch.repeat_interleave(torch.arange(paths.pair_count, device=key.device),
                                          paths.pair_offset[1:] - paths.pair_offset[:-1])
    assert torch.equal(paths.pair_tx[pair_of_row], tx_id)
    assert torch.equal(paths.pair_rx[pair_of_row], rx_id)
    pk = paths.pair_rx * t + paths.pair_tx
    assert bool((pk[1:] > pk[:-1]).all())
    comps = set(paths.rows["component_id"].unique().cpu().tolist())
    _say(f"components present: {sorted(comps)}")
    assert {0, 1, 2, 5} <= comps


def test_batched_paths_match_oracle_on_sampled_pairs(batch):
    w, compiled, dscene, inputs, paths = batch
    rows = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in paths.rows.items()}
    off = paths.pair_offset.cpu().numpy()
    pair_tx = paths.pair_tx.cpu().numpy()
    pair_rx = paths.pair_rx.cpu().numpy()
    rx_off = np.concatenate([[0], np.cumsum(w.rx_count)])
    orc = oracle.Oracle(compiled)
    rng = np.random.default_rng(33)
    txs = rng.choice(w.tx.shape[0], 6, replace=False)
    bits = component_bits(COMPONENTS)
    checked = 0
    worst_db = worst_ph = 0.0
    public = compiled.edge_public_id
    for t in txs:
        n = int(w.rx_count[t])
        local = np.sort(rng.choice(n, min(n, 16), replace=False))
        lo, hi = inputs.chains.offset_host[t * 5], inputs.chains.offset_host[t * 5 + 5]
        groups = inputs.chains.groups_host[lo:hi]
        depth = (groups >= 0).sum(axis=1).astype(np.int32)
        ref = orc.radiomap(tx=w.tx[t], rx=w.rx[int(rx_off[t]) + local], frequency=w.frequency, components=bits,
                           max_depth=3, chains=groups, chain_depth=depth, paths_cap=512, perturb=2e-4)
        for j, r in enumerate(local):
            if ref["ambiguous"][j]:
                continue
            pair = np.nonzero((pair_tx == t) & (pair_rx == r))[0]
            assert pair.size == 1
            a, b = int(off[pair[0]]), int(off[pair[0] + 1])
            ref_paths = ref["paths"][j]
            assert b - a == len(ref_paths), (t, r, b - a, len(ref_paths))
            for k, p in zip(range(a, b), ref_paths):
                assert rows["component_id"][k] == p.component and rows["depth"][k] == p.depth, (t, r, k)
                if p.component == 2:
                    assert rows["edge_id"][k] == public[p.edge]
                else:
                    for s in range(3):
                        assert rows["primitive_sequence"][k, s] == (p.prim[s] if s < p.depth else -1)
                assert abs(rows["path_length_m"][k] - p.length) <= 1e-5 * p.length + 1e-6
                if p.gain > 1e-25:
                    worst_db = max(worst_db, float(db_error(rows["path_gain"][k:k + 1], np.array([p.gain]))[0]))
                    coef = complex(p.coefficient[0], p.coefficient[1])
                    worst_ph = max(worst_ph, float(phase_error(rows["coefficient"][k:k + 1], np.array([coef]))[0]))
            checked += 1
    _say(f"checked {checked} pairs: worst path {worst_db:.2e} dB, worst coefficient phase {worst_ph:.2e} rad")
    assert checked >= 60
    assert worst_db <= DB_TOL
    assert worst_ph <= PHASE_TOL


def test_batched_paths_angles_and_delays(batch):
    w, compiled, dscene, inputs, paths = batch
    rows = paths.rows
    dep = rows["departure_direction"]
    arr = rows["arrival_direction"]
    st, ct = torch.sin(paths.theta_t), torch.cos(paths.theta_t)
    back = torch.stack([st * torch.cos(paths.phi_t), st * torch.sin(paths.phi_t), ct], dim=-1)
    # float32 acos loses about sqrt(ulp) near the poles: a wrong axis, sign or
    # convention would show as an O(1) difference
    assert float((back - dep / dep.norm(dim=-1, keepdim=True)).abs().max()) <= 1e-3
    sr, cr = torch.sin(paths.theta_r), torch.cos(paths.theta_r)
    back_r = torch.stack([sr * torch.cos(paths.phi_r), sr * torch.sin(paths.phi_r), cr], dim=-1)
    assert float((back_r + arr / arr.norm(dim=-1, keepdim=True)).abs().max()) <= 1e-3
    torch.testing.assert_close(paths.tau, rows["path_length_m"] / 299792458.0, rtol=1e-6, atol=0.0)


def test_batched_paths_cfr_matches_float64(batch):
    w, compiled, dscene, inputs, paths = batch
    freqs = torch.as_tensor(w.subcarriers)
    h = paths.cfr(freqs)
    torch.cuda.synchronize()
    assert h.shape == (paths.pair_count, 1024)
    rng = np.random.default_rng(7)
    off = paths.pair_offset.cpu().numpy()
    a = paths.a.cpu().numpy().astype(np.complex128)
    tau = paths.tau.cpu().numpy().astype(np.float64)
    hc = h.cpu().numpy()
    worst = 0.0
    for pr in rng.choice(paths.pair_count, 200, replace=False):
        lo, hi = int(off[pr]), int(off[pr + 1])
        ref = (a[lo:hi, None] * np.exp(-2j * math.pi * w.subcarriers[None, :] * tau[lo:hi, None])).sum(axis=0)
        scale = max(np.abs(a[lo:hi]).sum(), 1e-30)
        worst = max(worst, float(np.abs(hc[pr] - ref).max() / scale))
    _say(f"CFR: worst |error| / sum|a| over 200 pairs x 1024 subcarriers {worst:.2e}")
    assert worst <= 1e-5
