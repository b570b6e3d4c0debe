"""CPU restatement of the adjoint's work dealing (wc_vjp.hip vjp_kernel: per
tile, each lane's cell contributes its phase's descriptors; a wave prefix sum
flattens them, slice k of S takes iterations [k I / S, (k + 1) I / S), and
each path finds its owner lane by a binary-lifting search over the exclusive
prefix sums). Checked on random descriptor tables: every descriptor of the
phase is visited exactly once across the slices, and every index the kernel
forms (owner lane, descriptor slot, edge or chain id) stays in bounds."""

from __future__ import annotations

import numpy as np
import pytest

DIFF = 2


def _deal(counts_main, counts_diff, phase, slices):
    mine = np.where(phase == 0, counts_main, counts_diff).astype(np.int64)
    first = np.where(phase == 0, 0, counts_main).astype(np.int64)
    incl = np.cumsum(mine)
    excl = incl - mine
    total = int(incl[-1])
    iters = (total + 63) >> 6
    seen = []
    for k in range(slices):
        it0, it1 = iters * k // slices, iters * (k + 1) // slices
        for base in range(it0 * 64, it1 * 64, 64):
            for lane in range(64):
                j = base + lane
                u = 0
                for step in (32, 16, 8, 4, 2, 1):
                    probe = u + step
                    assert probe < 64
                    if excl[probe] <= j:
                        u = probe
                if j < total:
                    idx = first[u] + (j - excl[u])
                    assert first[u] <= idx < first[u] + mine[u]
                    seen.append((u, int(idx)))
    return seen, mine, first


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("phase,slices", [(0, 4), (0, 8), (1, 8), (1, 16)])
def test_dealing_visits_every_descriptor_once(seed, phase, slices):
    rng = np.random.default_rng(seed)
    live = rng.integers(1, 65)  # partial last tile
    counts_main = np.zeros(64, np.int64)
    counts_diff = np.zeros(64, np.int64)
    counts_main[:live] = rng.integers(0, 12, live)
    heavy = rng.uniform(size=live) < 0.2
    counts_diff[:live] = np.where(heavy, rng.integers(0, 400, live), rng.integers(0, 5, live))
    seen, mine, first = _deal(counts_main, counts_diff, phase, slices)
    want = {(u, int(first[u] + i)) for u in range(64) for i in range(int(mine[u]))}
    assert len(seen) == len(set(seen)) == len(want)
    assert set(seen) == want
