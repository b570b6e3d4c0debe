"""GPU tests of chain planning on the device (wc_plan.hip) against the host
restatement of the reference planner (reflection.py:186-205, 541-569: the
distinct prefixes of each transmitter's discovery rays per depth, sorted
lexicographically, or the exhaustive list for small group counts): the same
discovery rays give bit-identical chain tables, counts and CSR offsets, on
the urban canyon (exhaustive depth 1, discovered depths 2-3, 9 transmitters)
and on four transmitters of the scaled city (134,850 triangles, every depth
to 5 discovered)."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from witwin.channel import workloads
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _host_reference(compiled, rays: np.ndarray, max_depth: int) -> list[list[np.ndarray]]:
    g = compiled.group_count
    out = []
    for t in range(rays.shape[0]):
        per = []
        for d in range(1, engine.MAX_DEPTH + 1):
            if d > max_depth:
                per.append(np.zeros((0, d), np.int32))
            elif engine.is_exhaustive(g, d):
                per.append(engine.exhaustive_sequences(g, d))
            else:
                per.append(engine.unique_prefixes(rays[t], d))
        out.append(per)
    return out


def _check(compiled, dscene, tx: np.ndarray, max_depth: int, device):
    tx_t = torch.as_tensor(tx.astype(np.float32)).to(device)
    dseq = engine.chain_sequences_device(compiled, dscene, tx_t, max_depth, reflection=True)
    g = compiled.group_count
    needs = [d for d in range(1, max_depth + 1) if not engine.is_exhaustive(g, d)]
    rays = engine.discover(dscene, tx_t, max(needs), engine.DISCOVERY_RAY_COUNT) if needs else \
        np.zeros((tx.shape[0], 0, 0), np.int32)
    ref = _host_reference(compiled, rays, max_depth)
    torch.cuda.synchronize()
    assert dseq.discovered_depths == tuple(needs)
    for t in range(tx.shape[0]):
        got = dseq.host(t)
        for d in range(1, engine.MAX_DEPTH + 1):
            assert got[d - 1].shape == ref[t][d - 1].shape, (t, d, got[d - 1].shape, ref[t][d - 1].shape)
            assert np.array_equal(got[d - 1], ref[t][d - 1]), (t, d)
            assert dseq.counts[t, d - 1] == ref[t][d - 1].shape[0]
    # the table's transmitter and depth columns follow the CSR layout
    k = dseq.count
    tx_col = dseq.chain_tx[:k].cpu().numpy()
    depth_col = dseq.chain_depth[:k].cpu().numpy()
    off = dseq.offset_host
    for t in range(tx.shape[0]):
        for d in range(1, engine.MAX_DEPTH + 1):
            a, b = int(off[t * 5 + d - 1]), int(off[t * 5 + d])
            assert (tx_col[a:b] == t).all() and (depth_col[a:b] == d).all()
    # a plan built from the device table equals one built from host sequences
    plan_d = engine.plan_from_device(dscene, tx_t, dseq)
    plan_h = engine.plan_from_sequences(dscene, tx_t, ref, tuple(needs))
    torch.cuda.synchronize()
    assert np.array_equal(plan_d.offset_host, plan_h.offset_host)
    assert np.array_equal(plan_d.groups_host, plan_h.groups_host)
    assert torch.equal(plan_d.images[:k], plan_h.images[:k])
    return dseq


def test_device_planning_matches_host_on_the_canyon(device):
    _say("case: urban canyon seed 2, 9 transmitters, depth 3")
    scene = workloads.urban_canyon_scene(seed=2, shape=(8, 8), spacing=40.0)
    compiled = compile_scene(scene, reference_frequency_hz=28e9)
    dscene = device_scene(compiled, device)
    rng = np.random.default_rng(4)
    tx = np.column_stack([rng.uniform(0, 360, 9), rng.uniform(0, 360, 9), rng.uniform(3, 30, 9)])
    dseq = _check(compiled, dscene, tx, 3, device)
    _say(f"groups {compiled.group_count}, chains per transmitter {dseq.chain_counts.tolist()}")


def test_device_planning_matches_host_on_the_scaled_city(device):
    _say("case: scaled city seed 5, 4 transmitters, depth 5 (every depth discovered)")
    city = workloads.scaled_city(seed=5)
    compiled = compile_scene(city.scene, reference_frequency_hz=city.frequencies[-1])
    dscene = device_scene(compiled, device)
    dseq = _check(compiled, dscene, np.asarray(city.tx[:4]), 5, device)
    _say(f"groups {compiled.group_count}, chains per transmitter {dseq.chain_counts.tolist()}")
