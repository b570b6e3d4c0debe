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
