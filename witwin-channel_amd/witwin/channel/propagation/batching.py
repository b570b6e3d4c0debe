"""Many transmitters with their own receiver sets: batched radiomap forward +
fixed-topology VJP, plus the project autotuner that picks its launch shape.

The reference solves transmitters one at a time inside Python loops
(witwin/channel/propagation/enumerated.py:549-837, reflection.py:222-251,
diffraction.py:345-538). Here every transmitter of a batch shares one
persistent tiled launch per kernel (tiles never straddle transmitters), the
batches run back to back on one stream, and the whole sequence - every
batch at every carrier frequency, forward, loss gradient, adjoint and
deterministic reduction - is captured once in a HIP graph and replayed.

Launch shape (chosen by ``autotune`` unless given):

* ``visibility_mode``: chain and edge visibility by tile loops (receivers of
  a wavefront are neighbours) or by queued legs (scattered receivers);
* ``batch_size``: transmitters per launch. The chain-validity masks hold
  ``mask_words(max chains of one transmitter)`` words per cell, so a batch
  costs memory and mask traffic in proportion to its largest chain count;
  transmitters are ordered by chain count before batching so each batch is
  padded to a similar count. Larger batches amortize launches and tails.

The loss is L = mean over every lit (cell, carrier) of 10 log10(P); its
gradient with respect to the first-layer permittivity and conductivity of
every material, each carrier frequency and every transmitter position comes
from the same replay.
"""

from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import torch

from witwin.channel.propagation import engine, planning

AUTOTUNE_CACHE = Path(os.environ.get("WITWIN_AMD_CACHE", Path.home() / ".cache" / "witwin_channel_amd")) \
    / "autotune.json"
_LN10 = 2.302585092994046


@dataclass
class LaunchShape:
    visibility_mode: int          # 1 tile loops, 2 queued legs
    batch_size: int               # transmitters per launch
    source: str = "default"       # "autotuned", "cache" or "default"
    trials: dict[str, float] = field(default_factory=dict)  # ms per probe step of each candidate

    def as_dict(self) -> dict[str, Any]:
        return {"visibility_mode": {1: "tile_loops", 2: "queued_legs"}[self.visibility_mode],
                "batch_size": self.batch_size, "source": self.source, "trials_ms": self.trials}


def _pol(n: int) -> np.ndarray:
    return np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (n, 1))


class BatchedMeanDb:
    """Forward + VJP of the mean-dB loss over transmitters ``tx_pos`` [T, 3]
    whose receivers are the consecutive sets of sizes ``rx_count`` in
    ``rx_pos`` [sum, 3], at every carrier of ``frequencies``.

    ``launch()`` runs the step eagerly, ``capture()`` records it in a HIP
    graph and ``step()`` replays it. After a step: ``grad_params`` [2M]
    (dL/d eps_r, dL/d sigma per material, summed over carriers),
    ``grad_frequency`` [F] and ``grad_tx`` [T, 3] (transmitter order of the
    input)."""

    def __init__(self, compiled, dscene, tx_pos: np.ndarray, rx_pos: np.ndarray, rx_count: np.ndarray, *,
                 components: int, max_depth: int, frequencies: tuple[float, ...], shape: LaunchShape,
                 sequences: "planning.DeviceSequences | None" = None, tx_pol: np.ndarray | None = None,
                 rx_pol: np.ndarray | None = None, tx_power: np.ndarray | None = None,
                 coherent: bool = True) -> None:
        from witwin.channel.kernels.deterministic import MeanDbStep

        self.dscene = dscene
        self.device = dscene.device
        self.shape = shape
        t = int(tx_pos.shape[0])
        self.tx_count = t
        self.frequencies = tuple(float(f) for f in frequencies)
        self.material_count = compiled.materials.count
        rx_count = np.asarray(rx_count, dtype=np.int64)
        offsets = np.concatenate([[0], np.cumsum(rx_count)])
        tx_pol = _pol(t) if tx_pol is None else tx_pol
        rx_pol = _pol(rx_pos.shape[0]) if rx_pol is None else rx_pol
        tx_power = np.ones(t, np.float32) if tx_power is None else tx_power
        tx_t = torch.as_tensor(np.ascontiguousarray(tx_pos, dtype=np.float32)).to(self.device)
        t0 = time.perf_counter()
        if sequences is None:
            sequences = chain_sequences_batched(compiled, dscene, tx_t, max_depth, bool(components & 2))
        self.planning_s = time.perf_counter() - t0
        chains = np.asarray(sequences.chain_counts, dtype=np.int64)
        self.chain_counts = chains
        # transmitters with similar chain counts share a batch (the masks are
        # padded to the largest count of the batch)
        order = np.lexsort((np.arange(t), -chains))
        self.batches = [order[i:i + shape.batch_size] for i in range(0, t, max(shape.batch_size, 1))]
        self.units: list[tuple[int, np.ndarray, Any]] = []
        inputs_all = []
        for b, members in enumerate(self.batches):
            sl = [slice(int(offsets[i]), int(offsets[i + 1])) for i in members]
            rx_b = np.concatenate([rx_pos[s] for s in sl], axis=0)
            rxp_b = np.concatenate([rx_pol[s] for s in sl], axis=0)
            tx_b = torch.as_tensor(np.ascontiguousarray(tx_pos[members], dtype=np.float32)).to(self.device)
            plan = engine.plan_from_device(dscene, tx_b, sequences.subset(members))
            for fi, f in enumerate(self.frequencies):
                inputs_all.append((fi, members, engine.solve_inputs(
                    compiled, dscene, components=components, max_depth=max_depth, coherent=coherent,
                    frequency_hz=f, tx_pos=tx_b, tx_pol=tx_pol[members], tx_power=tx_power[members], rx_pos=rx_b,
                    rx_pol=rxp_b, rx_count=rx_count[members], chains=plan,
                    visibility_mode=shape.visibility_mode)))
        need = max(engine.workspace_bytes(dscene, inp) for _, _, inp in inputs_all)
        self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        for fi, members, inp in inputs_all:
            step = MeanDbStep(dscene, inp, self.material_count, workspace=self.workspace)
            if bool((step.out.failure != 0).any()):
                raise RuntimeError("batched radiomap exceeded a native capacity")
            self.units.append((fi, members, step))
        # one loss over every lit (cell, carrier): a shared normalization
        lit = sum(int((u[2].out.power_total > 0).sum().item()) for u in self.units)
        self.lit_cells = lit
        for _, _, step in self.units:
            power = step.out.power_total
            step.scale = torch.where(power > 0, torch.full_like(power, 10.0 / (_LN10 * max(lit, 1))),
                                     torch.zeros_like(power))
        self.cells = int(rx_count.sum())
        self.paths = sum(int(u[2].out.path_count.sum().item()) for u in self.units)
        m = self.material_count
        self.grad_params = torch.zeros(2 * m, dtype=torch.float32, device=self.device)
        self.grad_frequency = torch.zeros(len(self.frequencies), dtype=torch.float32, device=self.device)
        self.grad_tx = torch.zeros((t, 3), dtype=torch.float32, device=self.device)
        self._index = [torch.as_tensor(members.astype(np.int64)).to(self.device) for _, members, _ in self.units]
        self.graph: torch.cuda.CUDAGraph | None = None
        self.reuse_visibility = True

    @property
    def workspace_bytes(self) -> int:
        return int(self.workspace.numel())

    def launch(self) -> None:
        m = self.material_count
        self.grad_params.zero_()
        self.grad_frequency.zero_()
        self.grad_tx.zero_()
        for k, ((fi, _, step), idx) in enumerate(zip(self.units, self._index)):
            # the carriers of one batch run back to back: visibility masks are
            # geometry only, so every carrier after the first reuses them
            step.launch(reuse_visibility=self.reuse_visibility and fi > 0 and self.units[k - 1][0] == fi - 1)
            self.grad_params.add_(step.grad_params[:2 * m])
            self.grad_frequency[fi:fi + 1].add_(step.grad_params[2 * m:2 * m + 1])
            self.grad_tx.index_add_(0, idx, step.grad_tx)

    def capture(self) -> None:
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.launch()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.launch()
        torch.cuda.synchronize(self.device)

    def step(self) -> None:
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()

    def maps(self, frequency_index: int = 0, name: str = "power_total") -> list[torch.Tensor]:
        """Per-transmitter slices (input transmitter order) of one per-cell
        output of the radiomap (``power_total``, ``field_total``,
        ``path_count``) at one carrier."""
        maps: list[torch.Tensor | None] = [None] * self.tx_count
        for fi, members, step in self.units:
            if fi != frequency_index:
                continue
            values = getattr(step.out, name)
            off = step.inputs.cell_offset.cpu().numpy()
            for k, i in enumerate(members):
                maps[int(i)] = values[int(off[k]):int(off[k + 1])]
        return maps  # type: ignore[return-value]

    def power_maps(self, frequency_index: int = 0) -> list[torch.Tensor]:
        """Per-transmitter path-gain vectors (input transmitter order)."""
        return self.maps(frequency_index, "power_total")


def chain_sequences_batched(compiled, dscene, tx_pos: torch.Tensor, max_depth: int,
                            reflection: bool) -> "planning.DeviceSequences":
    """Candidate chains of every transmitter of ``tx_pos``, planned on the
    device in one pass (discovery rays of all transmitters, per-depth distinct
    prefixes, one chain table)."""
    return engine.chain_sequences_device(compiled, dscene, tx_pos, max_depth, reflection=reflection)


def _signature(compiled, tx_count: int, cells: int, components: int, max_depth: int, chains: int) -> str:
    return (f"faces={compiled.face_count},groups={compiled.group_count},edges={compiled.edge_count},tx={tx_count},"
            f"cells={cells},components={components},depth={max_depth},chains={chains}")


def _free_bytes(device) -> int:
    free, _total = torch.cuda.mem_get_info(device)
    return int(free)


def autotune(compiled, dscene, tx_pos: np.ndarray, rx_pos: np.ndarray, rx_count: np.ndarray, *, components: int,
             max_depth: int, frequencies: tuple[float, ...], sequences: "planning.DeviceSequences",
             memory_fraction: float = 0.45, use_cache: bool = True) -> LaunchShape:
    """Pick the launch shape of a batched solve by measurement, cached per
    workload signature.

    1. Visibility mode: both modes on one probe batch (the transmitters with
       the most chains, up to 8 of them), one forward + VJP step each after a
       warm-up; the faster wins.
    2. Batch size: the candidates 16, 32, 64 and all transmitters whose
       padded workspace fits ``memory_fraction`` of the free device memory;
       each is timed on the first 64 transmitters by chain count (all of
       them for the largest candidate) and the fastest per cell wins.
    """
    t = int(tx_pos.shape[0])
    rx_count = np.asarray(rx_count, dtype=np.int64)
    chains = np.asarray(sequences.chain_counts, dtype=np.int64)
    sig = _signature(compiled, t, int(rx_count.sum()), components, max_depth, int(chains.sum()))
    if use_cache and AUTOTUNE_CACHE.exists():
        try:
            hit = json.loads(AUTOTUNE_CACHE.read_text()).get(sig)
        except (OSError, ValueError):
            hit = None
        if hit:
            return LaunchShape(int(hit["visibility_mode"]), int(hit["batch_size"]), "cache",
                               dict(hit.get("trials", {})))
    offsets = np.concatenate([[0], np.cumsum(rx_count)])
    order = np.lexsort((np.arange(t), -chains))

    def subset(members: np.ndarray):
        sl = [slice(int(offsets[i]), int(offsets[i + 1])) for i in members]
        return (tx_pos[members], np.concatenate([rx_pos[s] for s in sl], axis=0), rx_count[members],
                sequences.subset(members))

    def trial(members: np.ndarray, shape: LaunchShape, reps: int = 2) -> float:
        txs, rxs, cnt, seqs = subset(members)
        run = BatchedMeanDb(compiled, dscene, txs, rxs, cnt, components=components, max_depth=max_depth,
                            frequencies=frequencies[:1], shape=shape, sequences=seqs)
        run.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run.launch()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        cells = run.cells
        del run
        torch.cuda.empty_cache()
        return ms / max(cells, 1)

    trials: dict[str, float] = {}
    probe = order[:min(8, t)]
    best_mode, best = 1, math.inf
    for mode in (1, 2):
        per_cell = trial(probe, LaunchShape(mode, len(probe)))
        trials[f"mode{mode}_probe{len(probe)}"] = per_cell * 1e3
        if per_cell < best:
            best_mode, best = mode, per_cell
    free = _free_bytes(dscene.device) * memory_fraction
    slice_members = order[:min(64, t)]
    best_b, best = min(16, t), math.inf
    for b in sorted({min(x, t) for x in (16, 32, 64, t)}):
        # padded workspace of the heaviest batch of this size
        heavy = order[:b]
        words = (int(chains[heavy].max()) + 31) // 32
        ws = int(rx_count[heavy].sum()) * (4 * words + 4 * ((compiled.edge_count + 31) // 32) + 32 + 4 * 256)
        if ws > free:
            continue
        members = slice_members if b <= len(slice_members) else order[:b]
        per_cell = trial(members, LaunchShape(best_mode, b))
        trials[f"batch{b}"] = per_cell * 1e3
        if per_cell < best:
            best_b, best = b, per_cell
    shape = LaunchShape(best_mode, best_b, "autotuned", trials)
    if use_cache:
        try:
            AUTOTUNE_CACHE.parent.mkdir(parents=True, exist_ok=True)
            cache = json.loads(AUTOTUNE_CACHE.read_text()) if AUTOTUNE_CACHE.exists() else {}
            cache[sig] = {"visibility_mode": shape.visibility_mode, "batch_size": shape.batch_size,
                          "trials": trials}
            AUTOTUNE_CACHE.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return shape
