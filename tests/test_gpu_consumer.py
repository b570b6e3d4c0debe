"""GPU tests of ``propagation.consumer.evaluate``: rows, pair segmentation and
transports against the deterministic export of the same endpoints, and the
line-of-sight Jones operator against its closed form."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from witwin.channel import deterministic, workloads
from witwin.channel.propagation import consumer
from witwin.channel.scene.compiler import compile as compile_scene
from witwin.core import Scene

pytestmark = pytest.mark.gpu

C0 = 299_792_458.0


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _batches(device, tx, tx_pow, rx, tx_pol=None, rx_pol=None, basis=False):
    nt, nr = len(tx), len(rx)
    tx_pol = np.tile([[0.0, 0.0, 1.0]], (nt, 1)) if tx_pol is None else tx_pol
    rx_pol = np.tile([[0.0, 0.0, 1.0]], (nr, 1)) if rx_pol is None else rx_pol
    f = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float32)).to(device).contiguous()  # noqa: E731
    b_t = f(np.tile([[[0, 0, 1], [1, 0, 0]]], (nt, 1, 1))) if basis else None
    b_r = f(np.tile([[[0, 0, 1], [0, 1, 0]]], (nr, 1, 1))) if basis else None
    sources = consumer.EndpointBatch(stable_ids=torch.arange(100, 100 + nt, device=device), positions_m=f(tx),
                                     polarizations=f(tx_pol), polarization_basis=b_t, powers_w=f(tx_pow))
    sinks = consumer.EndpointBatch(stable_ids=torch.arange(500, 500 + nr, device=device), positions_m=f(rx),
                                   polarizations=f(rx_pol), polarization_basis=b_r)
    return sources, sinks


def test_consumer_matches_deterministic_export(device):
    _say("case: consumer.evaluate urban canyon, 3 sources x 7 street sinks, depth 2, los+refl+diff")
    rng = np.random.default_rng(8)
    tx = [workloads.URBAN_CANYON_TX, (142.1, 97.3, 12.0), (301.7, 233.9, 6.5)]
    tx_pow = [1.0, 2.0, 0.5]
    rx = workloads.street_points(rng, 7)
    comps = frozenset({"los", "reflection", "diffraction"})
    base = workloads.urban_canyon_scene(seed=2, shape=(2, 2))
    compiled = compile_scene(base, reference_frequency_hz=28e9)
    sources, sinks = _batches(device, tx, tx_pow, rx)
    req = consumer.PropagationRequest(sources=sources, sinks=sinks, reference_frequency_hz=28e9, components=comps,
                                      max_depth=2, response="scalar_transport", topology_mode="discover",
                                      ad_mode="none")
    ev = consumer.evaluate(compiled, req)
    paths = ev.paths
    # the same endpoints as a scene: deterministic export rows, pair order rx-major
    scene = Scene(structures=list(base.structures),
                  endpoints=[workloads.transmitter(p, power_w=w) for p, w in zip(tx, tx_pow)]
                  + [workloads.receiver(tuple(map(float, r))) for r in rx])
    det = deterministic.solve(scene, deterministic.Config(max_depth=2, components=comps, export_paths=True),
                              reference_frequency_hz=28e9)
    rows = det.paths
    assert paths.path_count == int(rows.tx_id.shape[0]) > 0
    assert torch.equal(paths.topology.source_index.cpu(), rows.tx_id.cpu())
    assert torch.equal(paths.topology.sink_index.cpu(), rows.rx_id.cpu())
    assert torch.equal(paths.topology.primitive_sequence.cpu(), rows.primitive_sequence.cpu())
    coeff = torch.complex(rows.field_real, rows.field_imag).cpu()
    assert torch.equal(paths.transport.coefficient.cpu(), coeff)
    assert torch.equal(paths.geometry.delay_s.cpu(), rows.delay_s.cpu())
    # pair segmentation: pair = sink * 3 + source, offsets = exclusive scan of counts
    pair = (rows.rx_id.long() * 3 + rows.tx_id.long()).cpu()
    assert torch.equal(paths.pair_index.cpu(), pair)
    counts = torch.bincount(pair, minlength=21)
    assert paths.pair_offsets.cpu().tolist() == [0] + torch.cumsum(counts, 0).tolist()
    assert torch.equal(paths.topology.source_id.cpu(), 100 + rows.tx_id.long().cpu())
    assert torch.equal(paths.topology.sink_id.cpu(), 500 + rows.rx_id.long().cpu())
    # complex3: field x sqrt(power) projects back onto the scalar coefficient's magnitude
    req3 = consumer.PropagationRequest(sources=sources, sinks=sinks, reference_frequency_hz=28e9,
                                       components=comps, max_depth=2, response="complex3_transport",
                                       topology_mode="discover", ad_mode="none")
    ev3 = consumer.evaluate(compiled, req3)
    amp = torch.sqrt(torch.tensor(tx_pow, device=device))[paths.topology.source_index.long()]
    assert torch.allclose(ev3.paths.transport.field, rows.field_xyz * amp.unsqueeze(-1))
    capped = consumer.evaluate(compiled, consumer.PropagationRequest(
        sources=sources, sinks=sinks, reference_frequency_hz=28e9, components=comps, max_depth=2,
        response="scalar_transport", topology_mode="discover", ad_mode="none", max_paths=5))
    assert capped.paths.path_count == 5 and int(capped.paths.pair_offsets[-1]) == 5
    with pytest.raises(ValueError, match="does not match the compiled scene"):
        consumer.evaluate(compiled, consumer.PropagationRequest(
            sources=sources, sinks=sinks, reference_frequency_hz=3.5e9, components=comps, max_depth=2,
            response="scalar_transport", topology_mode="discover", ad_mode="none"))
    with pytest.raises(NotImplementedError, match="does not support components"):
        consumer.evaluate(compiled, consumer.PropagationRequest(
            sources=sources, sinks=sinks, reference_frequency_hz=28e9, components=comps, max_depth=2,
            response="polarimetric_transport", topology_mode="discover", ad_mode="none"))


def test_consumer_los_jones_closed_form(device):
    _say("case: consumer polarimetric_transport, free-space LoS, 2 sources x 3 sinks at 3.5 GHz")
    tx = [(0.0, 0.0, 10.0), (5.0, -3.0, 12.0)]
    rx = [(40.0, 0.0, 1.5), (0.0, 30.0, 1.5), (-20.0, -20.0, 3.0)]
    base = workloads.two_ray_scene(shape=(2, 2))
    compiled = compile_scene(base, reference_frequency_hz=3.5e9)
    sources, sinks = _batches(device, tx, [1.0, 4.0], rx, basis=True)
    req = consumer.PropagationRequest(sources=sources, sinks=sinks, reference_frequency_hz=3.5e9,
                                      components=frozenset({"los"}), max_depth=0,
                                      response="polarimetric_transport", topology_mode="discover",
                                      ad_mode="none")
    ev = consumer.evaluate(compiled, req)
    p = ev.paths
    assert p.path_count == 6
    j = p.transport
    k = 2 * np.pi * 3.5e9 / C0
    for row in range(6):
        s_i = int(p.topology.source_index[row])
        r_i = int(p.topology.sink_index[row])
        d = np.asarray(rx[r_i], dtype=np.float64) - np.asarray(tx[s_i], dtype=np.float64)
        dist = np.linalg.norm(d)
        carrier = np.exp(-1j * k * dist) / (2 * k * dist)
        tb = j.source_basis[row].cpu().numpy().astype(np.float64)
        rb = j.sink_basis[row].cpu().numpy().astype(np.float64)
        expect = np.array([[rb[i] @ tb[jj] for jj in range(2)] for i in range(2)]) * carrier
        got = j.matrix[row].cpu().numpy()
        assert np.allclose(got, expect, rtol=2e-4, atol=1e-6 * abs(carrier)), (row, got, expect)
        # bases transverse to the LoS direction
        dn = d / dist
        assert abs(tb[0] @ dn) < 1e-5 and abs(rb[1] @ dn) < 1e-5
