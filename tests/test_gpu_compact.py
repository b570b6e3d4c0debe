"""GPU tests of the device-side canonical compaction (wc_compact.hip) through
the C ABI, against the oracle's restatement of the reference canonical
compact (oracle.canonical_compact) and numpy: bit-exact permutations, keep
flags, kept offsets and gather lists on random row tables (empty pairs,
pairs of 1..700 rows so the LDS staging runs in several chunks, duplicate
identities, equal lengths, a NaN length, a per-pair cap); the canonical pair
order and row offsets against numpy's lexsort and cumsum."""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle
from witwin.channel import _native

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _random_rows(seed: int, width: int = 3):
    rng = np.random.default_rng(seed)
    sizes = np.concatenate([rng.integers(0, 6, 200), rng.integers(40, 80, 30), [0, 1, 257, 300, 700]])
    rng.shuffle(sizes)
    offset = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    k = int(offset[-1])
    depth = rng.integers(0, 4, k).astype(np.int32)
    comp = np.where(depth == 0, 0, rng.choice([1, 2, 5], k)).astype(np.int32)
    depth = np.where(comp == 2, 1, depth).astype(np.int32)
    seq = rng.integers(0, 3, (k, width)).astype(np.int32)
    seq[np.arange(width)[None, :] >= depth[:, None]] = -1
    prim = np.where(depth > 0, seq[:, 0], -1).astype(np.int32)
    edge = np.where(comp == 2, rng.integers(0, 4, k), -1).astype(np.int32)
    length = rng.choice(np.array([10.0, 10.5, 11.0, 12.25], np.float32), k).astype(np.float32)
    length[rng.integers(0, k)] = np.nan
    rows = {"depth": depth, "component_id": comp, "primitive_sequence": seq, "primitive_id": prim,
            "edge_id": edge, "path_length_m": length}
    return rows, offset


def _device_rows(rows, device):
    k = rows["depth"].shape[0]
    width = rows["primitive_sequence"].shape[1]
    t = {name: torch.as_tensor(np.ascontiguousarray(v)).to(device) for name, v in rows.items()}
    zeros_i = torch.zeros(k, dtype=torch.int32, device=device)
    t.update({"tx_id": zeros_i, "rx_id": zeros_i, "material_id": zeros_i})
    p = _native.ptr
    view = _native.WcPathRows(width=width, tx_id=p(t["tx_id"]), rx_id=p(t["rx_id"]), depth=p(t["depth"]),
                              component_id=p(t["component_id"]), primitive_id=p(t["primitive_id"]),
                              edge_id=p(t["edge_id"]), material_id=p(t["material_id"]),
                              path_length_m=p(t["path_length_m"]), primitive_sequence=p(t["primitive_sequence"]))
    return view, t


@pytest.mark.parametrize("cap", [None, 3])
@pytest.mark.parametrize("seed", [0, 1])
def test_device_compaction_matches_oracle(device, seed, cap):
    _say(f"case: random rows seed {seed}, cap {cap}")
    rows, offset = _random_rows(seed)
    ref = oracle.canonical_compact(rows, offset, cap=cap)
    view, keep_alive = _device_rows(rows, device)
    pairs = offset.shape[0] - 1
    k = int(offset[-1])
    p = _native.ptr
    pair_offset = torch.as_tensor(offset.astype(np.int32)).to(device)
    perm = torch.empty(k, dtype=torch.int32, device=device)
    keep = torch.empty(k, dtype=torch.uint8, device=device)
    kept_offset = torch.empty(pairs + 1, dtype=torch.int32, device=device)
    source = torch.full((k,), -1, dtype=torch.int32, device=device)
    total = torch.zeros(1, dtype=torch.int64, device=device)
    dropped = torch.zeros(1, dtype=torch.int32, device=device)
    ws = torch.empty(int(_native.call("wc_paths_compact_workspace_bytes", pairs)), dtype=torch.uint8, device=device)
    stream = _native.stream_handle(device)
    _native.call("wc_paths_compact", ctypes.byref(view), p(pair_offset), pairs, -1 if cap is None else cap, p(perm),
                 p(keep), p(kept_offset), p(source), p(total), p(dropped), p(ws), stream)
    torch.cuda.synchronize()
    n_kept = int(total.item())
    _say(f"{pairs} pairs, {k} rows, {n_kept} kept, {int(dropped.item())} dropped by the cap")
    assert np.array_equal(perm.cpu().numpy(), ref["perm"])
    assert np.array_equal(keep.cpu().numpy(), ref["keep"])
    assert np.array_equal(kept_offset.cpu().numpy(), ref["kept_offset"])
    assert n_kept == ref["source_row"].shape[0]
    assert np.array_equal(source[:n_kept].cpu().numpy(), ref["source_row"])
    assert int(dropped.item()) == ref["dropped"]
    # the order-only entry point gives the same permutation and flags
    perm2 = torch.empty_like(perm)
    keep2 = torch.empty_like(keep)
    _native.call("wc_paths_canonical_order", ctypes.byref(view), p(pair_offset), pairs, p(perm2), p(keep2), stream)
    torch.cuda.synchronize()
    assert torch.equal(perm2, perm) and torch.equal(keep2, keep)
    del keep_alive


@pytest.mark.parametrize("seed", [0, 1])
def test_pair_order_and_row_offsets_match_numpy(device, seed):
    rng = np.random.default_rng(100 + seed)
    t = 37
    rx_count = rng.integers(0, 3000, t)
    rx_count[rng.integers(0, t, 5)] = 0
    cells = int(rx_count.sum())
    _say(f"case: pair order of {t} transmitters, {cells} cells (seed {seed})")
    cell_offset = np.concatenate([[0], np.cumsum(rx_count)]).astype(np.int32)
    cell_tx = np.repeat(np.arange(t), rx_count)
    cell_rx = np.arange(cells) - cell_offset[cell_tx]
    want = np.lexsort((cell_tx, cell_rx))
    p = _native.ptr
    stream = _native.stream_handle(device)
    ws = torch.empty(int(_native.call("wc_paths_compact_workspace_bytes", cells)), dtype=torch.uint8, device=device)
    offs = torch.as_tensor(cell_offset).to(device)
    pair_cell = torch.empty(cells, dtype=torch.int32, device=device)
    _native.call("wc_paths_pair_order", p(offs), t, cells, int(rx_count.max()), p(pair_cell), p(ws), stream)
    counts = rng.integers(0, 40, cells).astype(np.int32)
    counts_d = torch.as_tensor(counts).to(device)
    row_offset = torch.empty(cells, dtype=torch.int32, device=device)
    pair_offset = torch.empty(cells + 1, dtype=torch.int32, device=device)
    total = torch.zeros(1, dtype=torch.int64, device=device)
    _native.call("wc_paths_row_offsets", p(counts_d), p(pair_cell), cells, p(row_offset), p(pair_offset), p(total),
                 p(ws), stream)
    torch.cuda.synchronize()
    assert np.array_equal(pair_cell.cpu().numpy(), want)
    scan = np.concatenate([[0], np.cumsum(counts[want])])
    assert np.array_equal(pair_offset.cpu().numpy(), scan)
    starts = np.empty(cells, np.int64)
    starts[want] = scan[:-1]
    assert np.array_equal(row_offset.cpu().numpy(), starts)
    assert int(total.item()) == int(counts.sum())
