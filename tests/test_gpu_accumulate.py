"""GPU tests of the flat accumulation facades (deterministic_accumulate_flat,
_fwd64, _backward, _jvp, _ad; csrc/wc_accum.hip) against the reference's
own fixture: its frozen SHA-256 of the six outputs for float32 and float64,
coherent and incoherent (tests/ad/test_deterministic_accum_ad.py:46-129),
bit for bit; strict float64 gradcheck of the adjoint and of the forward-mode
tangent; and float32 accumulation of random rows within float32 rounding of
the float64 oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from test_cpu_oracle_and_host import (ACCUM_CID, ACCUM_FIELDS, ACCUM_RX, ACCUM_SHA256, ACCUM_TX, accum_fixture,
                                      accum_sha256)
from witwin.channel.kernels import deterministic as ops

pytestmark = pytest.mark.gpu


def _rows(device, dtype):
    gain, real, imag = accum_fixture()
    t = lambda a, d: torch.as_tensor(np.asarray(a), dtype=d).to(device)  # noqa: E731
    valid = torch.ones(len(ACCUM_TX), dtype=torch.bool, device=device)
    return (valid, t(ACCUM_TX, torch.int32), t(ACCUM_RX, torch.int32), t(ACCUM_CID, torch.int32),
            t(gain, dtype), t(real, dtype), t(imag, dtype))


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("coherent", [True, False])
def test_accumulate_reproduces_reference_frozen_hash(device, dtype, coherent):
    print(f"case: reference accumulation fixture, {dtype}, coherent={coherent}", flush=True)
    rows = _rows(device, torch.float32 if dtype == "float32" else torch.float64)
    if dtype == "float32":
        out = ops.deterministic_accumulate_flat(*rows, num_tx=2, num_rx=2, coherent=coherent)
    else:
        out = ops.deterministic_accumulate_flat_ad(*rows, num_tx=2, num_rx=2, coherent=coherent)
    torch.cuda.synchronize()
    host = {k: out[k].detach().cpu().numpy() for k in ACCUM_FIELDS}
    assert host["power_total"].shape == (2, 2) and host["component_power"].shape == (6, 2, 2)
    assert accum_sha256(host) == ACCUM_SHA256[(dtype, coherent)]


@pytest.mark.parametrize("coherent", [True, False])
def test_accumulate_gradcheck_reverse_and_forward(device, coherent):
    print(f"case: accumulation gradcheck float64, coherent={coherent}", flush=True)
    valid, tx, rx, cid, gain, real, imag = _rows(device, torch.float64)
    leaves = tuple(x.clone().requires_grad_(True) for x in (gain, real, imag))

    def f(g, re, im):
        out = ops.deterministic_accumulate_flat_ad(valid, tx, rx, cid, g, re, im, num_tx=2, num_rx=2,
                                                   coherent=coherent)
        return tuple(out[k] for k in ACCUM_FIELDS)

    assert torch.autograd.gradcheck(f, leaves, eps=1e-6, atol=1e-8, rtol=1e-6, check_forward_ad=True,
                                    check_backward_ad=True, check_undefined_grad=False, check_batched_grad=False)


def test_accumulate_float32_matches_float64_oracle_on_random_rows(device):
    print("case: float32 accumulation vs the float64 oracle, 3000 random rows", flush=True)
    rng = np.random.default_rng(8)
    n, t, r = 3000, 3, 40
    tx = rng.integers(-1, t + 1, n).astype(np.int32)
    rx = rng.integers(0, r, n).astype(np.int32)
    cid = rng.choice([0, 1, 2, 5, 9], n).astype(np.int32)
    gain = rng.uniform(0, 1, n).astype(np.float32)
    re = rng.normal(size=n).astype(np.float32)
    im = rng.normal(size=n).astype(np.float32)
    valid = rng.uniform(size=n) > 0.1
    ref = oracle.accumulate_flat(valid, tx, rx, cid, gain, re, im, num_tx=t, num_rx=r, coherent=True,
                                 dtype=np.float64)
    d = lambda a, dt: torch.as_tensor(a, dtype=dt).to(device)  # noqa: E731
    out = ops.deterministic_accumulate_flat(d(valid, torch.bool), d(tx, torch.int32), d(rx, torch.int32),
                                            d(cid, torch.int32), d(gain, torch.float32), d(re, torch.float32),
                                            d(im, torch.float32), num_tx=t, num_rx=r, coherent=True)
    torch.cuda.synchronize()
    for k in ACCUM_FIELDS:
        np.testing.assert_allclose(out[k].cpu().numpy(), ref[k], rtol=2e-5, atol=2e-5)
