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


# ------------------------------------------------------------------ fixed-topology replay

def _replay_request(sources, sinks, topology, response="scalar_transport", **kw):
    return consumer.FixedTopologyRequest(sources=sources, sinks=sinks, reference_frequency_hz=28e9,
                                         topology=topology, response=response, ad_mode="none", **kw)


def _street_sinks(rng, n):
    """Sinks along the two streets that cross at the canyon transmitter, so
    most of them see it and its wall reflections."""
    along = rng.uniform(0.0, 370.0, n)
    across = rng.uniform(46.0, 59.0, n)
    ns = rng.uniform(size=n) < 0.5
    pts = np.stack([np.where(ns, across, along), np.where(ns, along, across), np.full(n, 1.5)], axis=1)
    return pts.astype(np.float32)


def _discover(device, tx, rx, compiled, comps=frozenset({"los", "reflection"}), depth=2):
    sources, sinks = _batches(device, tx, [1.0] * len(tx), rx)
    req = consumer.PropagationRequest(sources=sources, sinks=sinks, reference_frequency_hz=28e9, components=comps,
                                      max_depth=depth, response="scalar_transport", topology_mode="discover",
                                      ad_mode="none")
    return consumer.evaluate(compiled, req), sources, sinks


def _keys(paths, face_group):
    """(pair, component, coplanar groups of the sequence): a replayed row keeps
    its frozen face labels when a hit slides onto a coplanar twin triangle."""
    seq = paths.topology.primitive_sequence.cpu().numpy()
    groups = np.where(seq >= 0, face_group[np.maximum(seq, 0)], -1)
    comp = paths.topology.component_id.cpu().tolist()
    pair = paths.pair_index.cpu().tolist()
    return [(p, c, tuple(g)) for p, c, g in zip(pair, comp, groups.tolist())]


def test_replay_at_discovery_endpoints_reproduces_evaluate(device):
    _say("case: reevaluate at the discovery endpoints, canyon 2 sources x 9 sinks, depth 2, los+refl")
    rng = np.random.default_rng(21)
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    ev, sources, sinks = _discover(device, [workloads.URBAN_CANYON_TX, (52.7, 140.3, 12.0)],
                                   _street_sinks(rng, 9), compiled)
    prepared = consumer.prepare_fixed_topology(ev.paths.topology)
    out = consumer.reevaluate(compiled, _replay_request(sources, sinks, prepared))
    p, q = ev.paths, out.paths
    assert q.path_count == p.path_count > 0
    assert bool(out.row_valid.all())
    assert torch.equal(q.pair_index, p.pair_index) and torch.equal(q.pair_offsets, p.pair_offsets)
    a, b = q.transport.coefficient.cpu(), p.transport.coefficient.cpu()
    comp = p.topology.component_id.cpu()
    for c in (0, 1):
        sel = comp == c
        if bool(sel.any()):
            rel = ((a[sel] - b[sel]).abs() / b[sel].abs().clamp_min(1e-30)).max()
            _say(f"  component {c}: {int(sel.sum())} rows, identical {int((a[sel] == b[sel]).sum())}, "
                 f"max relative difference {float(rel):.2e}")
    assert torch.allclose(a, b, rtol=1e-6, atol=0.0)
    assert torch.equal(q.geometry.delay_s.cpu(), p.geometry.delay_s.cpu())
    assert torch.allclose(q.geometry.interaction_positions_m.cpu(), p.geometry.interaction_positions_m.cpu(),
                          rtol=0.0, atol=1e-5)
    assert torch.equal(q.geometry.interaction_normals.cpu(), p.geometry.interaction_normals.cpu())
    assert torch.allclose(q.geometry.field_direction.cpu(), p.geometry.field_direction.cpu(), rtol=0.0, atol=1e-6)
    assert out.diagnostics.validation_sync_count == 1 and out.diagnostics.discovery_launch_count == 0


def test_replay_with_moved_sinks_matches_fresh_discovery(device):
    _say("case: reevaluate with every sink moved 0.7 m vs a fresh evaluate at the moved sinks")
    rng = np.random.default_rng(22)
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    tx = [workloads.URBAN_CANYON_TX, (52.7, 140.3, 12.0)]
    rx = _street_sinks(rng, 12)
    ev, sources, _ = _discover(device, tx, rx, compiled)
    moved = rx + np.array([0.7, -0.4, 0.3], np.float32)
    _, sinks_moved = _batches(device, tx, [1.0] * 2, moved)
    out = consumer.reevaluate(compiled, _replay_request(sources, sinks_moved,
                                                        consumer.prepare_fixed_topology(ev.paths.topology)))
    fresh, _, _ = _discover(device, tx, moved, compiled)
    fresh_rows = {k: i for i, k in enumerate(_keys(fresh.paths, compiled.face_group))}
    valid = out.row_valid.cpu().tolist()
    coeff = out.paths.transport.coefficient.cpu()
    fcoeff = fresh.paths.transport.coefficient.cpu()
    kept = dropped = 0
    for r, key in enumerate(_keys(out.paths, compiled.face_group)):
        if valid[r]:
            assert key in fresh_rows, f"replayed row {key} is valid but absent from the fresh discovery"
            assert torch.allclose(coeff[r], fcoeff[fresh_rows[key]], rtol=1e-5, atol=1e-12)
            kept += 1
        else:
            assert key not in fresh_rows
            assert coeff[r] == 0
            dropped += 1
    _say(f"  rows kept {kept}, dropped {dropped}, new in the fresh discovery "
         f"{len(fresh_rows) - kept}")
    assert kept > 0


def test_replay_slots_and_wideband_columns(device):
    _say("case: reevaluate over 3 block-diagonal slots and 3 frequency offsets")
    rng = np.random.default_rng(23)
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    tx = [workloads.URBAN_CANYON_TX]
    rx = _street_sinks(rng, 5)
    ev, sources, sinks = _discover(device, tx, rx, compiled)
    prepared = consumer.prepare_fixed_topology(ev.paths.topology)
    shifts = [np.zeros(3, np.float32), np.array([0.3, 0.2, 0.0], np.float32), np.array([-0.5, 0.1, 0.2], np.float32)]
    singles = []
    for sh in shifts:
        _, s1 = _batches(device, tx, [1.0], rx + sh)
        singles.append(consumer.reevaluate(compiled, _replay_request(sources, s1, prepared)))
    src3, snk3 = _batches(device, tx * 3, [1.0] * 3, np.concatenate([rx + sh for sh in shifts]))
    rep = consumer.replicate_over_slots(prepared, 3, source_count=1, sink_count=5)
    out = consumer.reevaluate(compiled, _replay_request(src3, snk3, rep, slot_count=3))
    k = prepared.row_count
    assert out.paths.pair_count == 3 * 5 and out.paths.path_count == 3 * k
    for t, single in enumerate(singles):
        sl = slice(t * k, (t + 1) * k)
        assert torch.equal(out.paths.transport.coefficient[sl].cpu(), single.paths.transport.coefficient.cpu())
        assert torch.equal(out.row_valid[sl].cpu(), single.row_valid.cpu())
        assert torch.equal(out.paths.pair_index[sl].cpu() - t * 5, single.paths.pair_index.cpu())
    # wideband: column j equals a replay on the scene compiled at 28 GHz + df_j
    offsets = (-2.0e6, 0.0, 5.0e6)
    wb = consumer.reevaluate(compiled, _replay_request(sources, sinks, prepared, frequency_offsets_hz=offsets))
    cols = wb.paths.transport.coefficient_offsets
    assert cols.shape == (k, 3)
    assert torch.equal(cols[:, 1].cpu(), wb.paths.transport.coefficient.cpu())
    for j, df in enumerate(offsets):
        other = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9 + df)
        req = consumer.FixedTopologyRequest(sources=sources, sinks=sinks, reference_frequency_hz=28e9 + df,
                                            topology=prepared, response="scalar_transport", ad_mode="none")
        ref = consumer.reevaluate(other, req)
        assert torch.equal(cols[:, j].cpu(), ref.paths.transport.coefficient.cpu())


def _los_closed_form(src, snk, p_tx, p_rx, power, freq, response):
    """Free-space row in float64 torch (autograd through positions and
    frequency): E = (p_tx - (p_tx.k)k) / (2 k0 L) e^{-j k0 L}; the scalar
    coefficient projects E on p_rx - (p_rx.k)k and both carry sqrt(P)."""
    off = snk - src
    length = off.norm(dim=1)
    k = off / length[:, None]
    k0 = 2.0 * np.pi * freq / C0
    phase = torch.polar(torch.ones_like(length), -k0 * length) / (2.0 * k0 * length)
    e_perp = p_tx - (p_tx * k).sum(1, keepdim=True) * k
    amp = power.sqrt()
    if response == "scalar_transport":
        dip = (p_tx * p_rx).sum(1) - (p_tx * k).sum(1) * (p_rx * k).sum(1)
        return dip * phase * amp
    return e_perp * (phase * amp)[:, None]


@pytest.mark.parametrize("response", ["scalar_transport", "complex3_transport"])
def test_replay_los_rows_differentiate_endpoints_and_frequency(device, response):
    _say(f"case: reevaluate ad_mode vjp / jvp, line-of-sight rows, {response}, canyon 2 sources x 10 sinks")
    import torch.autograd.forward_ad as fwAD

    rng = np.random.default_rng(23)
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    tx = [workloads.URBAN_CANYON_TX, (52.7, 140.3, 12.0)]
    ev, sources, sinks = _discover(device, tx, _street_sinks(rng, 10), compiled, comps=frozenset({"los"}), depth=1)
    prepared = consumer.prepare_fixed_topology(ev.paths.topology)
    src_pos = sources.positions_m.detach().clone().requires_grad_(True)
    snk_pos = sinks.positions_m.detach().clone().requires_grad_(True)
    freq = torch.tensor(28e9, dtype=torch.float64, device=device, requires_grad=True)

    def run(ad_mode, s_pos, r_pos, f):
        s_b = consumer.EndpointBatch(stable_ids=sources.stable_ids, positions_m=s_pos,
                                     polarizations=sources.polarizations, powers_w=sources.powers_w)
        r_b = consumer.EndpointBatch(stable_ids=sinks.stable_ids, positions_m=r_pos, polarizations=sinks.polarizations)
        req = consumer.FixedTopologyRequest(sources=s_b, sinks=r_b, reference_frequency_hz=f, topology=prepared,
                                            response=response, ad_mode=ad_mode)
        out = consumer.reevaluate(compiled, req)
        t = out.paths.transport
        return out, (t.coefficient if response == "scalar_transport" else t.field)

    plain, values_plain = run("none", sources.positions_m, sinks.positions_m, 28e9)
    out, values = run("vjp", src_pos, snk_pos, freq)
    valid = out.row_valid
    assert valid is not None and int(valid.sum()) > 5
    torch.testing.assert_close(values.detach(), values_plain, rtol=2e-5, atol=1e-12)
    w = torch.as_tensor(rng.normal(size=tuple(values.shape) + (2,)), dtype=torch.float32, device=device)
    loss = (torch.view_as_real(values) * w).sum()
    loss.backward()
    # float64 closed form over the rows that still see each other
    topo = prepared.topology
    si, ri = topo.source_index.long(), topo.sink_index.long()
    s64 = sources.positions_m.detach().double().clone().requires_grad_(True)
    r64 = sinks.positions_m.detach().double().clone().requires_grad_(True)
    f64 = torch.tensor(28e9, dtype=torch.float64, device=device, requires_grad=True)
    ref = _los_closed_form(s64[si], r64[ri], sources.polarizations.double()[si], sinks.polarizations.double()[ri],
                           sources.powers_w.double()[si], f64, response)
    mask = valid.reshape((-1,) + (1,) * (ref.ndim - 1)).double()
    ref_loss = (torch.view_as_real(ref) * w.double() * mask[..., None]).sum()
    ref_loss.backward()
    for name, got, want in (("source", src_pos.grad, s64.grad), ("sink", snk_pos.grad, r64.grad),
                            ("frequency", freq.grad.reshape(1), f64.grad.reshape(1))):
        err = float((got.double() - want).norm() / want.norm().clamp_min(1e-300))
        _say(f"  d loss / d {name}: relative error {err:.2e} (|grad| {float(want.norm()):.3e})")
        assert err < 2e-3, name
    # forward mode: the tangent of the loss along a source-position direction
    direction = torch.as_tensor(rng.normal(size=tuple(src_pos.shape)), dtype=torch.float32, device=device)
    with fwAD.dual_level():
        dual = fwAD.make_dual(sources.positions_m.detach().clone(), direction)
        _, v = run("jvp", dual, sinks.positions_m, 28e9)
        tangent = fwAD.unpack_dual((torch.view_as_real(v) * w).sum()).tangent
    want = float((s64.grad * direction.double()).sum())
    assert abs(float(tangent) - want) <= 2e-3 * abs(want), (float(tangent), want)


def test_replay_reflection_rows_differentiate_endpoints(device):
    """Reflection rows under ad_mode vjp / jvp: values equal the plain replay;
    gradients with respect to source and sink positions and the frequency
    equal central differences of the FP64 oracle's rows (line of sight:
    oracle free_space; reflection: oracle reflection_sequence at interaction
    points re-solved by the image method over the frozen faces' planes)."""
    _say("case: reevaluate ad_mode vjp / jvp with reflection rows, canyon 2 sources x 8 sinks, depth 2")
    import torch.autograd.forward_ad as fwAD

    import oracle

    rng = np.random.default_rng(24)
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    tx = [workloads.URBAN_CANYON_TX, (52.7, 140.3, 12.0)]
    ev, sources, sinks = _discover(device, tx, _street_sinks(rng, 8), compiled)
    prepared = consumer.prepare_fixed_topology(ev.paths.topology)
    topo = prepared.topology

    def run(ad_mode, s_pos, r_pos, f):
        s_b = consumer.EndpointBatch(stable_ids=sources.stable_ids, positions_m=s_pos,
                                     polarizations=sources.polarizations, powers_w=sources.powers_w)
        r_b = consumer.EndpointBatch(stable_ids=sinks.stable_ids, positions_m=r_pos, polarizations=sinks.polarizations)
        req = consumer.FixedTopologyRequest(sources=s_b, sinks=r_b, reference_frequency_hz=f, topology=prepared,
                                            response="scalar_transport", ad_mode=ad_mode)
        return consumer.reevaluate(compiled, req)

    plain = run("none", sources.positions_m, sinks.positions_m, 28e9)
    src_pos = sources.positions_m.detach().clone().requires_grad_(True)
    snk_pos = sinks.positions_m.detach().clone().requires_grad_(True)
    freq = torch.tensor(28e9, dtype=torch.float64, device=device, requires_grad=True)
    out = run("vjp", src_pos, snk_pos, freq)
    coef = out.paths.transport.coefficient
    valid = out.row_valid.cpu().numpy()
    comp = topo.component_id.cpu().numpy()
    assert valid.sum() > 10 and (comp[valid] == 1).sum() > 5
    torch.testing.assert_close(coef.detach(), plain.paths.transport.coefficient, rtol=2e-5, atol=1e-12)
    w = rng.normal(size=(len(comp), 2)) * valid[:, None]
    loss = (torch.view_as_real(coef) * torch.as_tensor(w, dtype=torch.float32, device=device)).sum()
    loss.backward()

    # FP64 oracle loss over the valid rows
    si, ri = topo.source_index.cpu().numpy(), topo.sink_index.cpu().numpy()
    depth, seq = topo.depth.cpu().numpy(), topo.primitive_sequence.cpu().numpy()
    p_tx, p_rx = sources.polarizations.cpu().double().numpy(), sinks.polarizations.cpu().double().numpy()
    power = sources.powers_w.cpu().double().numpy()
    mats = compiled.materials

    def hits_of(a, b, planes):
        images, img = [], a
        for n, c in planes:
            img = img - 2.0 * (img @ n - c) * n
            images.append(img)
        out_h, nxt = [None] * len(planes), b
        for j in reversed(range(len(planes))):
            n, c = planes[j]
            s_i, s_n = images[j] @ n - c, nxt @ n - c
            out_h[j] = images[j] + (nxt - images[j]) * (s_i / (s_i - s_n))
            nxt = out_h[j]
        return out_h

    def oracle_loss(s_np, r_np, f):
        total = 0.0
        for row in np.nonzero(valid)[0]:
            a, b = s_np[si[row]], r_np[ri[row]]
            if comp[row] == 0:
                pf = oracle.free_space(a, b, power[si[row]], p_tx[si[row]], p_rx[ri[row]], f)["path_field"]
            else:
                faces = seq[row, :depth[row]]
                planes = [(compiled.group_plane[compiled.face_group[fc], :3].astype(np.float64),
                           float(compiled.group_plane[compiled.face_group[fc], 3])) for fc in faces]
                m = compiled.face_material[faces]
                pf = oracle.reflection_sequence(a, b, hits_of(a, b, planes), [n for n, _ in planes], power[si[row]],
                                                p_tx[si[row]], p_rx[ri[row]], mats.eps_r[m], mats.sigma_e[m],
                                                mats.mu_r[m], mats.gain[m], mats.thickness_m[m], f)["path_field"]
            total += w[row, 0] * pf.real + w[row, 1] * pf.imag
        return total

    s0 = sources.positions_m.cpu().double().numpy()
    r0 = sinks.positions_m.cpu().double().numpy()
    h = 1e-4

    def fd_grad(which):
        base = s0 if which == "source" else r0
        g = np.zeros_like(base)
        for idx in np.ndindex(*base.shape):
            plus, minus = base.copy(), base.copy()
            plus[idx] += h
            minus[idx] -= h
            args = (plus, r0) if which == "source" else (s0, plus)
            args_m = (minus, r0) if which == "source" else (s0, minus)
            g[idx] = (oracle_loss(*args, 28e9) - oracle_loss(*args_m, 28e9)) / (2 * h)
        return g

    g_src, g_snk = fd_grad("source"), fd_grad("sink")
    g_f = (oracle_loss(s0, r0, 28e9 + 1e3) - oracle_loss(s0, r0, 28e9 - 1e3)) / 2e3
    for name, got, want in (("source", src_pos.grad.cpu().double().numpy(), g_src),
                            ("sink", snk_pos.grad.cpu().double().numpy(), g_snk),
                            ("frequency", np.array([float(freq.grad)]), np.array([g_f]))):
        err = float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300))
        _say(f"  d loss / d {name}: relative error vs FP64 oracle central differences {err:.2e}")
        assert err < 2e-3, name
    direction = rng.normal(size=r0.shape)
    with fwAD.dual_level():
        dual = fwAD.make_dual(sinks.positions_m.detach().clone(),
                              torch.as_tensor(direction, dtype=torch.float32, device=device))
        v = run("jvp", sources.positions_m, dual, 28e9).paths.transport.coefficient
        tangent = float(fwAD.unpack_dual((torch.view_as_real(v) * torch.as_tensor(
            w, dtype=torch.float32, device=device)).sum()).tangent)
    want = float((g_snk * direction).sum())
    assert abs(tangent - want) <= 2e-3 * abs(want), (tangent, want)


def test_evaluate_ad_matches_replay_ad(device):
    """evaluate(ad_mode="vjp") discovers natively and differentiates the rows it
    found: values equal evaluate(ad_mode="none"), gradients equal those of the
    replay of the same rows (pinned against the FP64 oracle above); a request
    with diffraction is refused."""
    _say("case: evaluate ad_mode vjp, canyon 2 sources x 6 sinks, depth 2, los+refl, vs the replay's gradients")
    rng = np.random.default_rng(25)
    compiled = compile_scene(workloads.urban_canyon_scene(seed=2, shape=(2, 2)), reference_frequency_hz=28e9)
    tx = [workloads.URBAN_CANYON_TX, (52.7, 140.3, 12.0)]
    sources, sinks = _batches(device, tx, [1.0, 2.0], _street_sinks(rng, 6))
    comps = frozenset({"los", "reflection"})

    def batches(s_pos, r_pos):
        return (consumer.EndpointBatch(stable_ids=sources.stable_ids, positions_m=s_pos,
                                       polarizations=sources.polarizations, powers_w=sources.powers_w),
                consumer.EndpointBatch(stable_ids=sinks.stable_ids, positions_m=r_pos,
                                       polarizations=sinks.polarizations))

    def evaluate(ad_mode, s_pos, r_pos, components=comps):
        s_b, r_b = batches(s_pos, r_pos)
        return consumer.evaluate(compiled, consumer.PropagationRequest(
            sources=s_b, sinks=r_b, reference_frequency_hz=28e9, components=components, max_depth=2,
            response="scalar_transport", topology_mode="discover", ad_mode=ad_mode))

    plain = evaluate("none", sources.positions_m, sinks.positions_m)
    s1 = sources.positions_m.detach().clone().requires_grad_(True)
    r1 = sinks.positions_m.detach().clone().requires_grad_(True)
    ev = evaluate("vjp", s1, r1)
    coef = ev.paths.transport.coefficient
    assert ev.paths.path_count == plain.paths.path_count > 5
    torch.testing.assert_close(coef.detach(), plain.paths.transport.coefficient, rtol=2e-5, atol=1e-12)
    w = torch.as_tensor(rng.normal(size=(ev.paths.path_count, 2)), dtype=torch.float32, device=device)
    (torch.view_as_real(coef) * w).sum().backward()
    prepared = consumer.prepare_fixed_topology(plain.paths.topology)
    s2 = sources.positions_m.detach().clone().requires_grad_(True)
    r2 = sinks.positions_m.detach().clone().requires_grad_(True)
    s_b, r_b = batches(s2, r2)
    out = consumer.reevaluate(compiled, consumer.FixedTopologyRequest(
        sources=s_b, sinks=r_b, reference_frequency_hz=28e9, topology=prepared, response="scalar_transport",
        ad_mode="vjp"))
    assert bool(out.row_valid.all())
    (torch.view_as_real(out.paths.transport.coefficient) * w).sum().backward()
    torch.testing.assert_close(s1.grad, s2.grad, rtol=1e-5, atol=1e-9)
    torch.testing.assert_close(r1.grad, r2.grad, rtol=1e-5, atol=1e-9)
    with pytest.raises(NotImplementedError, match="line-of-sight and reflection"):
        evaluate("vjp", s1, r1, frozenset({"los", "reflection", "diffraction"}))
