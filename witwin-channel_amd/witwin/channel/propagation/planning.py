"""Chain planning on the device (reference witwin/channel/propagation/
reflection.py:186-205, 541-569): candidate reflection chains of many
transmitters as the distinct group-sequence prefixes of their discovery rays
(or the exhaustive sequence list where the group count is small), sorted per
(transmitter, depth) lexicographically, built without leaving the GPU: the
rays are sorted and de-duplicated depth by depth (wc_plan_level /
wc_plan_assign), the chain table is written by wc_plan_table /
wc_plan_exhaustive, and only the [T, 5] prefix counts come back to the host
(to size and lay out the table)."""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from witwin.channel import _native

MAX_DEPTH = 5


@dataclass
class DeviceSequences:
    """Candidate chains of T transmitters, rows ordered by (transmitter,
    depth, group sequence): ``groups`` [K, 5] (padded with -1), ``chain_tx``
    and ``chain_depth`` [K] on the device; ``counts`` [T, 5] and the CSR
    ``offset_host`` [T * 5 + 1] on the host (rows of (t, d) are
    offset_host[t*5 + d - 1] .. offset_host[t*5 + d])."""

    groups: torch.Tensor
    chain_tx: torch.Tensor
    chain_depth: torch.Tensor
    counts: np.ndarray
    offset_host: np.ndarray
    discovered_depths: tuple[int, ...] = ()
    _host: dict[int, list[np.ndarray]] = field(default_factory=dict, repr=False)

    def __len__(self) -> int:
        return int(self.counts.shape[0])

    def __getitem__(self, t: int) -> list[np.ndarray]:
        return self.host(int(t))

    @property
    def chain_counts(self) -> np.ndarray:
        """Chains per transmitter."""
        return self.counts.sum(axis=1)

    @property
    def count(self) -> int:
        return int(self.offset_host[-1])

    def host(self, t: int) -> list[np.ndarray]:
        """Transmitter t's sequences per depth 1..5 as host arrays [K_d, d]."""
        if t not in self._host:
            lo, hi = int(self.offset_host[t * MAX_DEPTH]), int(self.offset_host[t * MAX_DEPTH + MAX_DEPTH])
            rows = self.groups[lo:hi].cpu().numpy() if hi > lo else np.zeros((0, MAX_DEPTH), np.int32)
            per = []
            for d in range(1, MAX_DEPTH + 1):
                a = int(self.offset_host[t * MAX_DEPTH + d - 1]) - lo
                b = int(self.offset_host[t * MAX_DEPTH + d]) - lo
                per.append(np.ascontiguousarray(rows[a:b, :d]).astype(np.int32))
            self._host[t] = per
        return self._host[t]

    def subset(self, members) -> "DeviceSequences":
        """The sequences of transmitters ``members`` (in that order, renumbered
        0..len-1)."""
        members = np.asarray(members, dtype=np.int64)
        device = self.groups.device
        counts = self.counts[members]
        offset = np.concatenate([[0], np.cumsum(counts.reshape(-1))]).astype(np.int64)
        spans = [np.arange(int(self.offset_host[m * MAX_DEPTH]), int(self.offset_host[m * MAX_DEPTH + MAX_DEPTH]),
                           dtype=np.int64) for m in members]
        rows = np.concatenate(spans) if spans else np.zeros(0, np.int64)
        k = rows.shape[0]
        if k:
            idx = torch.as_tensor(rows).to(device)
            groups = self.groups.index_select(0, idx)
            depth = self.chain_depth.index_select(0, idx)
        else:
            groups = torch.full((1, MAX_DEPTH), -1, dtype=torch.int32, device=device)
            depth = torch.zeros(1, dtype=torch.int32, device=device)
        local = np.repeat(np.arange(len(members), dtype=np.int32), counts.sum(axis=1)) if k else np.zeros(1, np.int32)
        return DeviceSequences(groups=groups, chain_tx=torch.as_tensor(local).to(device), chain_depth=depth,
                               counts=counts, offset_host=offset, discovered_depths=self.discovered_depths)


def exhaustive_sequences(group_count: int, depth: int) -> np.ndarray:
    """All adjacent-distinct group sequences of one depth, lexicographic."""
    if group_count <= 0:
        return np.zeros((0, depth), np.int32)
    seqs = np.arange(group_count, dtype=np.int32)[:, None]
    for _ in range(depth - 1):
        nxt = np.arange(group_count, dtype=np.int32)
        a = np.repeat(seqs, group_count, axis=0)
        b = np.tile(nxt, seqs.shape[0])[:, None]
        keep = a[:, -1] != b[:, 0]
        seqs = np.concatenate([a[keep], b[keep]], axis=1)
    return seqs


def device_sequences(group_count: int, dscene, tx_pos: torch.Tensor, max_depth: int, *, reflection: bool,
                     exhaustive, ray_count: int) -> DeviceSequences:
    """Candidate sequences of every transmitter of ``tx_pos`` [T, 3]: for each
    depth up to ``max_depth`` the exhaustive list when ``exhaustive(d)``,
    else the distinct d-prefixes of the transmitter's ``ray_count``
    discovery rays."""
    device = tx_pos.device
    stream = _native.stream_handle(device)
    p = _native.ptr
    t = int(tx_pos.shape[0])
    g = int(group_count)
    depths = list(range(1, max_depth + 1)) if reflection and g > 0 and t > 0 else []
    needs_rays = [d for d in depths if not exhaustive(d)]
    counts = np.zeros((t, MAX_DEPTH), np.int64)
    levels = []
    if needs_rays:
        dmax = max(needs_rays)
        rays = torch.empty((t, ray_count, dmax), dtype=torch.int32, device=device)
        _native.call("wc_discover_chains", dscene.ref(), p(tx_pos), t, ray_count, dmax, p(rays), stream)
        ws = torch.empty(int(_native.call("wc_plan_workspace_bytes", t, ray_count)), dtype=torch.uint8,
                         device=device)
        counts_d = torch.zeros((t, MAX_DEPTH), dtype=torch.int32, device=device)
        total = torch.zeros(1, dtype=torch.int64, device=device)
        parent_tx = None
        for d in range(1, dmax + 1):
            _native.call("wc_plan_level", p(rays), t, ray_count, dmax, d, g, p(counts_d), p(total), p(ws), stream)
            k = int(total.item())
            arrays = [torch.empty(max(k, 1), dtype=torch.int32, device=device) for _ in range(3)]
            _native.call("wc_plan_assign", p(rays), t, ray_count, dmax, d, g, p(parent_tx), p(arrays[0]),
                         p(arrays[1]), p(arrays[2]), p(counts_d), p(ws), stream)
            levels.append((k, arrays))
            parent_tx = arrays[2]
        discovered = counts_d.cpu().numpy().astype(np.int64)
        del rays, ws
    tables = {}
    for d in depths:
        if d in needs_rays:
            counts[:, d - 1] = discovered[:, d - 1]
        else:
            tables[d] = exhaustive_sequences(g, d)
            counts[:, d - 1] = tables[d].shape[0]
    offset = np.concatenate([[0], np.cumsum(counts.reshape(-1))]).astype(np.int64)
    k_all = int(offset[-1])
    if k_all >= 2 ** 31:
        raise RuntimeError(f"chain planning: {k_all} chains exceed the int32 chain index")
    groups = torch.full((max(k_all, 1), MAX_DEPTH), -1, dtype=torch.int32, device=device)
    chain_tx = torch.zeros(max(k_all, 1), dtype=torch.int32, device=device)
    chain_depth = torch.zeros(max(k_all, 1), dtype=torch.int32, device=device)
    if levels:
        view = _native.WcPlanLevels()
        for d, (k, arrays) in enumerate(levels, start=1):
            view.prefix_parent[d - 1] = p(arrays[0])
            view.prefix_group[d - 1] = p(arrays[1])
            view.prefix_tx[d - 1] = p(arrays[2])
    for d in depths:
        block = offset[np.arange(t) * MAX_DEPTH + d - 1]
        if d in needs_rays:
            first_id = np.concatenate([[0], np.cumsum(counts[:, d - 1])[:-1]])
            row_start = torch.as_tensor((block - first_id).astype(np.int32)).to(device)
            _native.call("wc_plan_table", ctypes.byref(view), d, int(levels[d - 1][0]), p(row_start), p(groups),
                         p(chain_tx), p(chain_depth), stream)
        elif tables[d].shape[0]:
            seqs = torch.as_tensor(np.ascontiguousarray(tables[d], dtype=np.int32)).to(device)
            row_start = torch.as_tensor(block.astype(np.int32)).to(device)
            _native.call("wc_plan_exhaustive", p(seqs), int(tables[d].shape[0]), d, t, p(row_start), p(groups),
                         p(chain_tx), p(chain_depth), stream)
    return DeviceSequences(groups=groups, chain_tx=chain_tx, chain_depth=chain_depth, counts=counts,
                           offset_host=offset, discovered_depths=tuple(needs_rays))
