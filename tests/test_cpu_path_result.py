"""Host-side logic of the path API (ragged storage, padding, CIR, taps,
type filtering, configuration checks) on CPU tensors. The expected values
are worked out by hand from the reference contract (witwin/channel/path.py:
pair order :274-297, padding fills :586-635, cir :637-644, taps :658-699,
filter_by_type :734-788)."""

from __future__ import annotations

import math

import pytest
import torch

from witwin.channel.path import (
    Config,
    InteractionType,
    PathResult,
    RaggedPathSoA,
    endpoint_angles,
    interaction_types,
)


def _ragged(max_paths_per_pair=None) -> RaggedPathSoA:
    # three paths over 2 receivers x 1 transmitter, given out of pair order:
    # rx1 LoS (100 ns), rx1 reflection (150 ns), rx0 reflection (200 ns)
    rx = torch.tensor([1, 1, 0])
    tx = torch.tensor([0, 0, 0])
    field = torch.tensor([1 + 0j, 0.5j, -0.25 + 0j], dtype=torch.complex64)
    delay = torch.tensor([100e-9, 150e-9, 200e-9], dtype=torch.float32)
    itype = torch.tensor([[0], [1], [1]], dtype=torch.int32)
    prim = torch.tensor([[-1], [7], [3]], dtype=torch.int32)
    mat = torch.tensor([[-1], [0], [1]], dtype=torch.int32)
    pos = torch.arange(9, dtype=torch.float32).reshape(3, 1, 3)
    zeros = torch.zeros(3)
    return RaggedPathSoA.from_flat(num_rx=2, num_rx_ant=1, num_tx=1, num_tx_ant=1, rx_id=rx, tx_id=tx, field=field,
                                   delay_s=delay, theta_t=zeros, phi_t=zeros, theta_r=zeros, phi_r=zeros,
                                   interaction_type=itype, primitive_id=prim, material_id=mat, position=pos,
                                   normal=pos.clone(), max_paths_per_pair=max_paths_per_pair)


def test_ragged_orders_pairs_and_keeps_input_order_inside_a_pair():
    r = _ragged()
    assert r.pair_offsets.tolist() == [0, 1, 3]
    assert r.rx_id.tolist() == [0, 1, 1]
    assert r.delay_s.tolist() == pytest.approx([200e-9, 100e-9, 150e-9])
    assert r.primitive_id[:, 0].tolist() == [3, -1, 7]
    r1 = _ragged(max_paths_per_pair=1)
    assert r1.pair_offsets.tolist() == [0, 1, 2]
    assert r1.delay_s.tolist() == pytest.approx([200e-9, 100e-9])


def test_ragged_rejects_bad_shapes_and_ids():
    r = _ragged()
    with pytest.raises(ValueError, match="rx_id is outside"):
        RaggedPathSoA.from_flat(num_rx=1, num_rx_ant=1, num_tx=1, num_tx_ant=1, rx_id=r.rx_id, tx_id=r.tx_id,
                                field=r.field, delay_s=r.delay_s, theta_t=r.theta_t, phi_t=r.phi_t,
                                theta_r=r.theta_r, phi_r=r.phi_r, interaction_type=r.interaction_type,
                                primitive_id=r.primitive_id, material_id=r.material_id, position=r.position,
                                normal=r.normal)
    with pytest.raises(ValueError, match="field must use complex64"):
        RaggedPathSoA.from_flat(num_rx=2, num_rx_ant=1, num_tx=1, num_tx_ant=1, rx_id=r.rx_id, tx_id=r.tx_id,
                                field=r.field.to(torch.complex128), delay_s=r.delay_s, theta_t=r.theta_t,
                                phi_t=r.phi_t, theta_r=r.theta_r, phi_r=r.phi_r,
                                interaction_type=r.interaction_type, primitive_id=r.primitive_id,
                                material_id=r.material_id, position=r.position, normal=r.normal)


def test_padding_fills_and_path_counts():
    res = PathResult.from_ragged(_ragged())
    assert res.a.shape == (2, 1, 1, 1, 2, 1)
    assert res.num_paths.reshape(-1).tolist() == [1, 2]
    assert res.valid.reshape(2, 2).tolist() == [[True, False], [True, True]]
    assert res.tau.reshape(2, 2)[0, 1].item() == -1.0
    assert res.primitive_id.reshape(2, 2)[0, 1].item() == -1
    assert res.material_id.reshape(2, 2)[0, 1].item() == -1
    assert res.interaction_type.reshape(2, 2)[0, 1].item() == int(InteractionType.NONE)
    assert res.a.reshape(2, 2)[1, 1].item() == 0.5j
    assert res.path_length_m.reshape(2, 2)[1, 0].item() == pytest.approx(100e-9 * 299_792_458.0, rel=1e-6)
    assert res.path_length_m.reshape(2, 2)[0, 1].item() == -1.0
    assert res.metadata["schema"] == "PathResult"
    wide = PathResult.from_ragged(_ragged(), minimum_path_width=4)
    assert wide.max_num_paths == 4


def test_cir_normalizes_to_first_arrival_per_pair():
    res = PathResult.from_ragged(_ragged())
    a, tau = res.cir()
    t = tau.reshape(2, 2)
    assert t[0].tolist() == pytest.approx([0.0, -1.0])
    assert t[1].tolist() == pytest.approx([0.0, 50e-9], rel=1e-6, abs=1e-15)
    _, raw = res.cir(normalize_delays=False)
    assert raw.reshape(2, 2)[1].tolist() == pytest.approx([100e-9, 150e-9])
    assert a.reshape(2, 2)[0, 1].item() == 0


def test_taps_bin_by_rounded_delay():
    res = PathResult.from_ragged(_ragged())
    taps = res.taps(1e8, 8).reshape(2, 8)
    # rx0: one path at tap 0; rx1: LoS at tap 0, reflection 50 ns -> tap 5
    assert taps[0].tolist() == [-0.25 + 0j] + [0j] * 7
    expected = [0j] * 8
    expected[0] = 1 + 0j
    expected[5] = 0.5j
    assert taps[1].tolist() == expected
    with pytest.raises(ValueError, match="bandwidth"):
        res.taps(0.0, 4)


def test_filter_by_type_keeps_matching_paths():
    res = PathResult.from_ragged(_ragged())
    refl = res.filter_by_type(int(InteractionType.REFLECTION))
    assert refl.num_paths.reshape(-1).tolist() == [1, 1]
    assert refl.tau.reshape(2, -1)[:, 0].tolist() == pytest.approx([200e-9, 150e-9])
    los = res.filter_by_type(0)
    assert los.num_paths.reshape(-1).tolist() == [0, 1]
    assert los.metadata["filtered_interaction_types"] == [0]


def test_interaction_type_slots():
    comp = torch.tensor([0, 1, 2, 5, 1])
    depth = torch.tensor([0, 2, 1, 3, 1])
    t = interaction_types(comp, depth, 3)
    assert t.tolist() == [[0, 0, 0], [1, 1, 0], [2, 0, 0], [4, 4, 4], [1, 0, 0]]


def test_endpoint_angles_closed_form():
    d = torch.tensor([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, -1.0]])
    theta, phi = endpoint_angles(d)
    assert theta.tolist() == pytest.approx([0.0, math.pi / 2, 3 * math.pi / 4], abs=1e-6)
    assert phi.tolist() == pytest.approx([0.0, 0.0, math.pi / 2], abs=1e-6)


def test_path_config_validation():
    assert Config(max_depth=3).components == frozenset({"los", "reflection", "diffraction"})
    with pytest.raises(ValueError, match="per_pair"):
        Config(max_paths_scope="global")
    with pytest.raises(ValueError, match="max_paths must be positive"):
        Config(max_paths=0)
    with pytest.raises(RuntimeError, match="max_depth <= 5"):
        Config(max_depth=6, components={"reflection"})
    with pytest.raises(ValueError, match="ad_mode"):
        Config(ad_mode="forward")
