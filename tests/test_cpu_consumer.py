"""Host-side logic of the propagation consumer: endpoint validation, the
vocabulary and capability record, and the transverse-basis construction of
the Jones response (orthonormal, transverse, oriented)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from witwin.channel.propagation import consumer


def test_endpoint_batch_rejects_host_tensors_and_bad_dtypes():
    with pytest.raises(ValueError, match="CUDA tensors"):
        consumer.EndpointBatch(stable_ids=torch.zeros(2, dtype=torch.int64), positions_m=torch.zeros((2, 3)),
                               polarizations=torch.zeros((2, 3)))
    with pytest.raises(TypeError, match="positions_m must have dtype"):
        consumer.EndpointBatch(stable_ids=torch.zeros(2, dtype=torch.int64),
                               positions_m=torch.zeros((2, 3), dtype=torch.float64),
                               polarizations=torch.zeros((2, 3)))
    with pytest.raises(ValueError, match=r"shape \(N, 3\)"):
        consumer.EndpointBatch(stable_ids=torch.zeros(2, dtype=torch.int64), positions_m=torch.zeros((2, 2)),
                               polarizations=torch.zeros((2, 3)))


def test_capabilities_vocabulary():
    caps = consumer.capabilities()
    assert caps.contract_version == 6
    assert caps.components_for("polarimetric_transport") == frozenset({"los"})
    assert caps.components_for("scalar_transport") == consumer.COMPONENTS
    assert consumer.RESPONSES == {"scalar_transport", "complex3_transport", "polarimetric_transport"}
    conv = consumer.PropagationConvention()
    assert conv.pair_layout.endswith("pair_index=sink_index*source_count+source_index")


def test_transverse_basis_is_orthonormal_transverse_and_oriented():
    rng = np.random.default_rng(3)
    d = torch.as_tensor(rng.normal(size=(64, 3)), dtype=torch.float32)
    ref = torch.as_tensor(rng.normal(size=(64, 2, 3)), dtype=torch.float32)
    ref[0, 0] = d[0] * 2.5            # first reference vector along the direction: falls back to the second
    b = consumer.transverse_basis(ref, d)
    u, v = b[:, 0], b[:, 1]
    dn = d / d.norm(dim=-1, keepdim=True)
    assert torch.allclose((u * u).sum(-1), torch.ones(64), atol=1e-5)
    assert torch.allclose((v * v).sum(-1), torch.ones(64), atol=1e-5)
    assert torch.allclose((u * v).sum(-1), torch.zeros(64), atol=1e-5)
    assert torch.allclose((u * dn).sum(-1), torch.zeros(64), atol=1e-5)
    assert torch.allclose((v * dn).sum(-1), torch.zeros(64), atol=1e-5)
    # u follows the first reference vector, v points along the second
    assert bool(((u[1:] * ref[1:, 0]).sum(-1) > 0).all())
    assert bool(((v * ref[:, 1]).sum(-1) >= 0).all())
    w = ref[0, 1] - (ref[0, 1] @ dn[0]) * dn[0]
    assert torch.allclose(u[0], w / w.norm(), atol=1e-5)


def _topology(component, depth, seq, source=None, sink=None):
    k = len(component)
    i32 = lambda v: torch.as_tensor(v, dtype=torch.int32)  # noqa: E731
    source = i32(source if source is not None else [0] * k)
    sink = i32(sink if sink is not None else list(range(k)))
    seq = i32(seq).reshape(k, -1)
    return consumer.PropagationTopology(
        source_index=source, sink_index=sink, source_id=source.long() + 100, sink_id=sink.long() + 500,
        depth=i32(depth), component_id=i32(component), primitive_id=seq[:, 0].clone() if seq.shape[1] else i32([-1] * k),
        edge_id=i32([-1] * k), material_id=i32([0] * k), primitive_sequence=seq, material_sequence=seq.clone(),
        interaction_type=seq.clone())


def test_prepare_fixed_topology_buckets_keep_row_order():
    topo = _topology([0, 1, 1, 0, 1], [0, 2, 1, 0, 1], [[-1, -1], [4, 7], [3, -1], [-1, -1], [9, -1]])
    prepared = consumer.prepare_fixed_topology(topo)
    got = [(b.component, b.depth, b.rows.tolist()) for b in prepared.buckets]
    # ascending (component, depth) keys, ascending rows inside a bucket
    assert got == [("los", 0, [0, 3]), ("reflection", 1, [2, 4]), ("reflection", 2, [1])]
    assert prepared.row_count == 5 and prepared.slot_count == 1


def test_prepare_fixed_topology_rejects_padding_and_components():
    with pytest.raises(ValueError, match="padding"):
        consumer.prepare_fixed_topology(_topology([1], [1], [[4, 5]]))
    with pytest.raises(ValueError, match="padding"):
        consumer.prepare_fixed_topology(_topology([1], [2], [[4, -1]]))
    with pytest.raises(NotImplementedError, match="component id 2"):
        consumer.prepare_fixed_topology(_topology([2], [1], [[11, -1]]))
    with pytest.raises(TypeError):
        consumer.prepare_fixed_topology("rows")


def test_replicate_over_slots_is_index_arithmetic():
    topo = _topology([0, 1], [0, 1], [[-1], [6]], source=[1, 0], sink=[0, 2])
    prepared = consumer.prepare_fixed_topology(topo)
    assert consumer.replicate_over_slots(prepared, 1, source_count=2, sink_count=3) is prepared
    rep = consumer.replicate_over_slots(prepared, 3, source_count=2, sink_count=3)
    t = rep.topology
    assert rep.slot_count == 3 and t.row_count == 6
    assert t.source_index.tolist() == [1, 0, 3, 2, 5, 4]
    assert t.sink_index.tolist() == [0, 2, 3, 5, 6, 8]
    assert t.primitive_sequence[:, 0].tolist() == [-1, 6] * 3
    assert [(b.component, b.rows.tolist()) for b in rep.buckets] == [("los", [0, 2, 4]), ("reflection", [1, 3, 5])]
    with pytest.raises(ValueError, match="already replicated"):
        consumer.replicate_over_slots(rep, 2, source_count=2, sink_count=3)
    with pytest.raises(ValueError, match="slot_count"):
        consumer.replicate_over_slots(prepared, 0, source_count=2, sink_count=3)


def test_frequency_offsets_validation_and_capabilities():
    assert consumer._require_frequency_offsets(None) is None
    assert consumer._require_frequency_offsets((0, -1e6, 2.5e6)) == (0.0, -1e6, 2.5e6)
    for bad in ((), [1.0], (float("nan"),), (True,)):
        with pytest.raises(ValueError):
            consumer._require_frequency_offsets(bad)
    caps = consumer.capabilities()
    assert caps.supports_fixed_topology and caps.fixed_topology_components == {"los", "reflection"}
    assert caps.wideband_responses == {"scalar_transport", "complex3_transport"}
    assert consumer.PropagationConvention().slot_pair_layout.startswith("block_diagonal_slots:")
