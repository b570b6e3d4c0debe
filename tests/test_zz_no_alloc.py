"""Allocation-free flat accumulation against the workspace entry point.

``wc_accumulate_flat_noalloc`` keeps its scratch in a fixed device array; it
must give bit-identical outputs to ``wc_accumulate_flat`` on the same rows,
also on a second call with a smaller input (the scratch then holds the first
call's contents), and refuse inputs whose scratch need exceeds its size.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from witwin.channel import _native

pytestmark = pytest.mark.gpu


def _rows(seed: int, count: int, num_tx: int, num_rx: int, device):
    rng = np.random.default_rng(seed)
    comp = rng.choice(np.array([0, 1, 2, 5, 6, 3], np.int32), size=count)
    rows = {
        "valid": torch.as_tensor((rng.uniform(size=count) < 0.95).astype(np.uint8)),
        "tx": torch.as_tensor(rng.integers(0, num_tx, size=count, dtype=np.int32)),
        "rx": torch.as_tensor(rng.integers(0, num_rx, size=count, dtype=np.int32)),
        "comp": torch.as_tensor(comp),
        "gain": torch.as_tensor(rng.uniform(1e-9, 1e-5, size=count).astype(np.float32)),
        "re": torch.as_tensor(rng.normal(size=count).astype(np.float32) * 1e-3),
        "im": torch.as_tensor(rng.normal(size=count).astype(np.float32) * 1e-3),
    }
    return {k: v.to(device) for k, v in rows.items()}


def _outputs(cells: int, device):
    return (torch.full((6, cells), -1.0, dtype=torch.float32, device=device),
            torch.full((6, cells, 2), -1.0, dtype=torch.float32, device=device),
            torch.full((cells,), -1.0, dtype=torch.float32, device=device),
            torch.full((cells, 2), -1.0, dtype=torch.float32, device=device))


def _run(noalloc: bool, rows, count, num_tx, num_rx, coherent, device):
    p = _native.ptr
    outs = _outputs(num_tx * num_rx, device)
    args = [count] + [p(rows[k]) for k in ("valid", "tx", "rx", "comp", "gain", "re", "im")] + \
        [num_tx, num_rx, coherent] + [p(o) for o in outs]
    stream = _native.stream_handle(device)
    if noalloc:
        _native.call("wc_accumulate_flat_noalloc", *args, stream)
    else:
        ws = torch.empty(int(_native.call("wc_accumulate_workspace_bytes", count, num_tx, num_rx)),
                         dtype=torch.uint8, device=device)
        _native.call("wc_accumulate_flat", *args, p(ws), ws.numel(), stream)
    torch.cuda.synchronize(device)
    return [o.cpu() for o in outs]


@pytest.mark.parametrize("coherent", [1, 0])
def test_noalloc_accumulate_matches_workspace_entry_point(device, coherent):
    for call, (count, num_tx, num_rx, seed) in enumerate(((200_000, 7, 3_000, 11), (5_000, 3, 400, 12))):
        print(f"case: accumulate noalloc call {call + 1}, {count} rows, {num_tx}x{num_rx} cells, "
              f"coherent={coherent}", flush=True)
        sys.stdout.flush()
        rows = _rows(seed, count, num_tx, num_rx, device)
        ref = _run(False, rows, count, num_tx, num_rx, coherent, device)
        got = _run(True, rows, count, num_tx, num_rx, coherent, device)
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
        assert float(ref[2].abs().sum()) > 0.0


def test_noalloc_accumulate_refuses_oversized_inputs(device):
    limit = int(_native.call("wc_accumulate_noalloc_max_bytes"))
    count, num_tx, num_rx = 2_000_000, 4, 4_000_000
    assert int(_native.call("wc_accumulate_workspace_bytes", count, num_tx, num_rx)) > limit
    dummy = torch.zeros(1, dtype=torch.int32, device=device)
    p = _native.ptr(dummy)
    with pytest.raises(RuntimeError, match="scratch"):
        _native.call("wc_accumulate_flat_noalloc", count, *([p] * 7), num_tx, num_rx, 1, *([p] * 4),
                     _native.stream_handle(device))
