"""Radiomap forward+VJP throughput on one MI355X.

Workload of the value line: the scaled city (seed 5): 400 random boxes on a
256 x 256 fBm heightfield (134,850 triangles), 8 materials, 128 transmitters
3-30 m above the terrain, each with its own square receiver grid of side
U{64..512} spanning the terrain (12,417,953 cells in all), 5-bounce specular
reflection + first-order UTD diffraction + thin-sheet transmission + line of
sight, at 3.5 GHz and 28 GHz. One step = for every transmitter and both
carriers the radiomap forward (path descriptors stored) + the fixed-topology
VJP of L = mean dB path gain over every lit (cell, carrier) w.r.t. every
material's eps_r and sigma, both carrier frequencies and every transmitter
position, with inputs resident in HBM. ``value`` = cells x carriers / step
time of the batched form (transmitter batches sized by the autotuner, one
persistent tiled launch per kernel and batch, the whole step captured once
in a HIP graph and replayed). The plain form is the per-transmitter,
per-carrier loop of the same kernels, launched eagerly.

Also reported: the urban canyon (1 Tx, 256 x 256 grid), a 64-transmitter
point-set batch on it, the differentiable heightfield with vertex gradients,
the roofline of the dominant kernel and the FP64 CPU oracle on a bounded
sample of the scaled city.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "witwin-channel_amd"), str(ROOT)]

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes of each launch: what it must read and write at least
# once. Per cell: receiver position (12 B) + evaluation-order entry (4 B);
# the reflection pass writes three component slots (power f32 + field c64 =
# 12 B each) and the 32 B float64 partial sums handed to the diffraction pass;
# the diffraction pass reads those 32 B and writes its slot (12 B), the totals
# (power + field, 12 B), the diffraction vector (3 x c64, 24 B) and the path
# count (4 B); the adjoint reads the cell's dL/dP (4 B), totals (12 B), four
# component fields (32 B), diffraction vector (24 B) and path count (4 B).
# Each path's 4 B descriptor is written by the pass that finds it and read by
# the adjoint. Scene tables (triangles 48 B, BVH nodes 64 B, edges 80 B, chain
# groups + float64 images 140 B) count once per launch.
BYTES_CELL_MAIN = 12 + 4 + 36 + 32
BYTES_CELL_DIFF = 12 + 4 + 32 + 12 + 12 + 24 + 4
BYTES_CELL_VJP = 12 + 4 + 4 + 12 + 32 + 24 + 4
BYTES_PER_DESCRIPTOR = 4


def timed(fn, steps: int, warmup: int, stream) -> tuple[float, list[float]]:
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    wall0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    return wall / steps * 1e3, [s.elapsed_time(e) for s, e in zip(starts, ends)]


def kernel_time(fn, stream, reps: int = 5, before=None) -> float:
    """Average duration (ms) of ``fn``'s launches, bracketed by HIP events on
    the stream they run on; ``before`` (untimed) runs ahead of each rep."""
    import torch

    if before is not None:
        before()
    fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    for i in range(reps):
        if before is not None:
            before()
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in zip(starts, ends)) / reps


def traffic_bytes(kernel: str, workload: str):
    """Measured HBM bytes per launch of ``kernel`` on ``workload`` from the
    committed PMC summary (profiles/hbm_traffic.json: FETCH_SIZE + WRITE_SIZE
    scaled by a copy kernel), or None when no summary matches."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return None
    entry = json.loads(path.read_text()).get("workloads", {}).get(workload, {}).get("kernels", {}).get(kernel)
    return entry.get("bytes_per_launch") if entry else None


def descriptor_split(out) -> tuple[int, int]:
    """(reflection-pass paths, diffraction paths) stored in the topology table."""
    import torch

    cap = out.topology_capacity
    cols = torch.arange(cap, device=out.topology.device).unsqueeze(0)
    valid = cols < out.path_count.unsqueeze(1)
    diff = valid & ((out.topology >> 24) == 2)
    total = int(valid.sum().item())
    nd = int(diff.sum().item())
    return total - nd, nd


def copy_calibration(nbytes: int = 256 << 20, reps: int = 3) -> int:
    """A plain device-to-device copy of ``nbytes`` (reads and writes nbytes
    each), the known quantity that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}))
    return 0


def scaled_city_section(steps: int, warmup: int, stream, cpu: bool, plain_loop: bool = True,
                        launch_shape: str | None = None) -> dict:
    """The value workload: batched fwd+VJP over the scaled city (see the
    module docstring); returns the measurements for the JSON line."""
    import numpy as np
    import torch

    from witwin.channel import workloads
    from witwin.channel.components import component_bits
    from witwin.channel.propagation import batching, engine
    from witwin.channel.scene.compiler import compile as compile_scene, device_scene

    device = torch.device("cuda", 0)
    depth = 5
    bits = component_bits({"los", "reflection", "diffraction", "transmission"})
    city = workloads.scaled_city(seed=5)
    compiled = compile_scene(city.scene, reference_frequency_hz=city.frequencies[-1])
    dscene = device_scene(compiled, device)
    tx_t = torch.as_tensor(city.tx).to(device)
    t0 = time.perf_counter()
    seqs = batching.chain_sequences_batched(compiled, dscene, tx_t, depth, True)
    torch.cuda.synchronize()
    planning_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    if launch_shape:
        mode, size = (int(x) for x in launch_shape.split(","))
        shape = batching.LaunchShape(mode, size, "given")
    else:
        shape = batching.autotune(compiled, dscene, city.tx, city.rx, city.rx_count, components=bits,
                                  max_depth=depth, frequencies=city.frequencies, sequences=seqs)
    tune_s = time.perf_counter() - t0
    run = batching.BatchedMeanDb(compiled, dscene, city.tx, city.rx, city.rx_count, components=bits,
                                 max_depth=depth, frequencies=city.frequencies, shape=shape, sequences=seqs)
    run.capture()
    ms_batched, per_step = timed(run.step, steps, warmup, stream)
    cells = run.cells * len(city.frequencies)
    grads = {"d_eps_r": run.grad_params[:run.material_count].cpu().tolist(),
             "d_sigma": run.grad_params[run.material_count:].cpu().tolist(),
             "d_frequency": run.grad_frequency.cpu().tolist(),
             "d_tx_norm_mean": float(run.grad_tx.norm(dim=1).mean())}

    # kernels of the heaviest unit (first batch, first carrier), each alone
    # with HIP events on its stream
    _, members, unit = run.units[0]
    inp = unit.inputs
    p = lambda k: (lambda: engine.radiomap_pass(dscene, inp, unit.out, k))  # noqa: E731

    def before(k):
        return lambda: [engine.radiomap_pass(dscene, inp, unit.out, j) for j in range(1, k)]

    unit_ms = None
    if plain_loop:
        # (skipped with --value-only, so a kernel trace of that run averages every batch evenly)
        unit_ms = {"chain_validity": kernel_time(p(1), stream),
                   "fields_main": kernel_time(p(2), stream, before=before(2)),
                   "edge_visibility": kernel_time(p(3), stream, before=before(3)),
                   "fields_diffraction": kernel_time(p(4), stream, before=before(4))}
        unit.launch()
        unit_ms["adjoint_main"] = kernel_time(lambda: unit.adjoint(1), stream)
        unit_ms["adjoint_diffraction"] = kernel_time(lambda: unit.adjoint(2), stream, before=lambda: unit.adjoint(1))
    unit_cells = inp.total_cells
    scene_bytes = compiled.face_count * 48 + int(dscene.node_count) * 64
    # the dominant kernel, the chain validity pass, over every batch (first
    # carrier; the second reuses its masks): per cell it reads the receiver
    # and its order entry and writes one bit per chain of its own
    # transmitter; the scene and the chain tables (groups + float64 images)
    # once per launch
    first = [u for u in run.units if u[0] == 0]
    tx_chains = np.asarray(seqs.chain_counts, np.int64)
    cell_bytes = city.rx_count.astype(np.int64) * (16 + 4 * ((tx_chains + 31) // 32))

    def all_chain_passes():
        for _, _, u in first:
            engine.radiomap_pass(dscene, u.inputs, u.out, 1)

    dom_ms = kernel_time(all_chain_passes, stream, reps=3) / len(first)
    dom_bytes = sum(int(cell_bytes[np.asarray(m)].sum()) + scene_bytes + int(u.inputs.chains.count) * (4 * 5 + 8 * 15)
                    for _, m, u in first) / len(first)
    paths = run.paths

    ms_plain, identical = None, None
    if plain_loop:
        plain = batching.BatchedMeanDb(compiled, dscene, city.tx, city.rx, city.rx_count, components=bits,
                                       max_depth=depth, frequencies=city.frequencies,
                                       shape=batching.LaunchShape(shape.visibility_mode, 1), sequences=seqs)
        plain.reuse_visibility = False
        ms_plain, _ = timed(plain.launch, 2, 1, stream)
        identical = all(torch.equal(a, b) for fi in range(len(city.frequencies))
                        for a, b in zip(run.power_maps(fi), plain.power_maps(fi)))
        del plain

    cpu_line = None
    if cpu:
        from tests.gpu_helpers import DB_TOL
        import oracle

        t = 0
        rng = np.random.default_rng(0)
        n = int(city.rx_count[t])
        sample = np.sort(rng.choice(n, 64, replace=False))
        per = seqs.host(t)
        groups = np.concatenate([np.pad(x, ((0, 0), (0, 5 - x.shape[1])), constant_values=-1) for x in per if x.size],
                                axis=0).astype(np.int32)
        chain_depth = (groups >= 0).sum(axis=1).astype(np.int32)
        orc = oracle.Oracle(compiled)
        c0 = time.perf_counter()
        ref = orc.radiomap(tx=city.tx[t], rx=city.rx[:n][sample], frequency=city.frequencies[-1], components=bits,
                           max_depth=depth, chains=groups, chain_depth=chain_depth, perturb=2e-4)
        cpu_s = time.perf_counter() - c0
        p_gpu = run.power_maps(len(city.frequencies) - 1)[t].cpu().numpy()[sample].astype(np.float64)
        p_ref = ref["record"][:, 0]
        amb = ref["ambiguous"].astype(bool)
        lit = (p_gpu > 0) & (p_ref > 0)
        err = np.abs(10 * np.log10(p_gpu[lit] / p_ref[lit]))
        stable = lit & ~amb
        err_s = np.abs(10 * np.log10(p_gpu[stable] / p_ref[stable]))
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu_line = {
            "value": len(sample) * 5 / cpu_s, "unit": "cells/s (forward only, each cell evaluated 5 times: at its "
                                                    "receiver and with receiver/transmitter moved by 0.2 mm)",
            "cores": threads, "kind": "port",
            "sample": f"64 random cells of transmitter 0's {int(city.sides[t])}x{int(city.sides[t])} grid at 28 GHz "
                      f"(its {groups.shape[0]} candidate chains, 1600 edges, 134,850 triangles brute force), "
                      f"FP64 oracle, OpenMP over cells, {cpu_s:.1f} s",
            "cells_lit_both": int(lit.sum()), "cells_lit_one_only": int(((p_gpu > 0) != (p_ref > 0)).sum()),
            "ambiguous_cells": int(amb.sum()),
            "median_db_error_vs_gpu": float(np.median(err)) if err.size else 0.0,
            "max_db_error_vs_gpu": float(err.max()) if err.size else 0.0,
            "cells_above_0.01_db": int((err > DB_TOL).sum()),
            "max_db_error_topology_stable_cells": float(err_s.max()) if err_s.size else 0.0,
            "stable_cells_above_0.01_db": int((err_s > DB_TOL).sum()),
        }

    return {
        "cells": cells, "ms_batched": ms_batched, "ms_plain": ms_plain, "per_step_ms": per_step,
        "planning_s": planning_s, "tune_s": tune_s, "shape": shape.as_dict(), "paths": paths,
        "identical": identical, "grads": grads, "unit_ms": unit_ms, "unit_cells": unit_cells,
        "unit_tx": int(len(members)), "dom_ms": dom_ms, "dom_bytes": dom_bytes, "cpu": cpu_line,
        "tx": int(city.tx.shape[0]), "grid_cells": int(run.cells), "triangles": compiled.face_count,
        "edges": compiled.edge_count, "chains": int(seqs.count),
        "workspace_gb": run.workspace_bytes / 1e9, "frequencies": list(city.frequencies),
        "batches": len(run.batches),
    }


def public_solve_section(grid: int, frequency: float, depth: int, reps: int = 3) -> dict:
    """The public call on the urban canyon, end to end: ``deterministic.solve``
    with ``ad_mode="vjp"`` (scene compile or cache lookup, chain planning,
    inputs, forward) and ``backward()`` of the mean-dB loss to the transmitter
    position, wall time with a synchronization at the end. The first call
    compiles the scene; the later ones find it in the compile cache."""
    import torch

    from witwin.channel import workloads
    from witwin.channel.deterministic import Config, solve

    scene = workloads.urban_canyon_scene(seed=2, shape=(grid, grid), spacing=2.0)
    tx = scene.endpoints[0]
    tx.position.requires_grad_(True)
    cfg = Config(components={"los", "reflection", "diffraction"}, max_depth=depth, ad_mode="vjp")

    def once() -> float:
        tx.position.grad = None
        t0 = time.perf_counter()
        res = solve(scene, cfg, reference_frequency_hz=frequency)
        p = res.path_gain
        lit = p > 0
        loss = (10.0 * torch.log10(torch.where(lit, p, torch.ones_like(p)))).sum() / lit.sum().clamp_min(1)
        loss.backward()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    cold = once()
    warm = sorted(once() for _ in range(reps))
    return {"call": "deterministic.solve(ad_mode='vjp') + backward() to the transmitter position",
            "cold_ms": cold, "warm_ms": warm[len(warm) // 2], "warm_ms_all": warm,
            "grad_tx": [float(x) for x in tx.position.grad]}


def batched_paths_section(stream) -> dict:
    """Batched explicit-path solve (urban canyon + 3 layered street walls, 64
    transmitters with 16..4096 street points each, LoS / reflection /
    transmission to depth 3 + UTD, 28 GHz) and its channel frequency response
    on 1,024 subcarriers over 100 MHz."""
    import numpy as np
    import torch

    from witwin.channel import workloads
    from witwin.channel.components import component_bits
    from witwin.channel.path import Config, paths_cfr, solve_batched
    from witwin.channel.propagation import engine
    from witwin.channel.scene.compiler import compile as compile_scene, device_scene

    device = torch.device("cuda", 0)
    comps = {"los", "reflection", "diffraction", "transmission"}
    w = workloads.batched_path_workload(seed=3, tx_count=64)
    compiled = compile_scene(w.scene, reference_frequency_hz=w.frequency)
    dscene = device_scene(compiled, device)
    bits = component_bits(comps)
    t = w.tx.shape[0]
    pol = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (w.rx.shape[0], 1))
    t0 = time.perf_counter()
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=3, coherent=True,
                                 frequency_hz=w.frequency, tx_pos=w.tx, tx_pol=pol[:t],
                                 tx_power=np.ones(t, np.float32), rx_pos=w.rx, rx_pol=pol, rx_count=w.rx_count)
    torch.cuda.synchronize()
    planning_s = time.perf_counter() - t0
    cfg = Config(components=comps, max_depth=3)
    paths = solve_batched(compiled, cfg, w.tx, w.rx, w.rx_count, dscene=dscene, inputs=inputs)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        paths = solve_batched(compiled, cfg, w.tx, w.rx, w.rx_count, dscene=dscene, inputs=inputs)
    torch.cuda.synchronize()
    solve_ms = (time.perf_counter() - t0) * 1e3 / reps
    a, tau, valid = paths.padded()
    freqs = torch.as_tensor(w.subcarriers)
    cfr_ms = kernel_time(lambda: paths_cfr(a, tau, valid, freqs, normalize_delays=False), stream)
    pairs, width = int(a.shape[0]), int(a.shape[1])
    cfr_bytes = pairs * 1024 * 8 + pairs * width * (8 + 4 + 1)
    return {
        "workload": f"urban canyon + 3 layered street walls ({compiled.face_count} tris), {t} Tx with street point "
                    f"sets 16..4096 ({pairs} pairs), LoS/reflection/transmission depth 3 + UTD, 28 GHz; CFR on "
                    "1024 subcarriers over 100 MHz",
        "pairs": pairs, "paths": paths.path_count, "chain_planning_s": planning_s,
        "solve_ms": solve_ms, "paths_per_s": paths.path_count / (solve_ms * 1e-3),
        "cfr_ms": cfr_ms, "cfr_padded_paths_per_pair": width,
        "cfr_roofline": {"bound": "hbm", "achieved_gbs": cfr_bytes / (cfr_ms * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
                         "frac": cfr_bytes / (cfr_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes": cfr_bytes},
    }


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--canyon-steps", type=int, default=20)
    parser.add_argument("--canyon-only", action="store_true",
                        help="skip the value workload; print the urban canyon measurements only")
    parser.add_argument("--value-only", action="store_true",
                        help="only the value workload (no plain loop, canyon, heightfield or CPU oracle)")
    parser.add_argument("--launch-shape", default=None,
                        help="visibility_mode,batch_size for the value workload instead of the autotuner")
    parser.add_argument("--grid", type=int, default=256)
    parser.add_argument("--tx-batch", type=int, default=64)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--no-batch", action="store_true")
    parser.add_argument("--no-heightfield", action="store_true")
    parser.add_argument("--copy-calibration", action="store_true",
                        help="only run a plain device copy of known size (HBM counter calibration)")
    args = parser.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py measures one GPU; runs on several GPUs are not part of this benchmark "
              "(the radiomap shards by transmitter and receiver tile with no communication)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from witwin.channel import workloads
    from witwin.channel.components import component_bits
    from witwin.channel.kernels.deterministic import MeanDbStep
    from witwin.channel.propagation import engine
    from witwin.channel.scene.compiler import DeviceScene, compile as compile_scene, device_scene, with_vertices

    device = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(device)
    city = None
    if not args.canyon_only:
        city = scaled_city_section(args.steps, args.warmup, stream, cpu=not args.no_cpu and not args.value_only,
                                   plain_loop=not args.value_only, launch_shape=args.launch_shape)
        torch.cuda.empty_cache()
    if args.value_only:
        print(json.dumps({"value_only": True, "cells_per_s": city["cells"] / (city["ms_batched"] * 1e-3),
                          "ms_per_step": city["ms_batched"], "launch_shape": city["shape"],
                          "chain_pass_ms_per_batch": city["dom_ms"], "heaviest_batch_ms": city["unit_ms"]}))
        return 0
    frequency = 28.0e9
    depth = 3
    bits = component_bits({"los", "reflection", "diffraction"})
    scene = workloads.urban_canyon_scene(seed=2, shape=(args.grid, args.grid), spacing=2.0)
    compiled = compile_scene(scene, reference_frequency_hz=frequency)
    dscene = device_scene(compiled, device)
    t0 = time.perf_counter()
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                 frequency_hz=frequency)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    cells = inputs.total_cells
    m = compiled.materials.count

    step = MeanDbStep(dscene, inputs, m)
    if bool((step.out.failure != 0).any()):
        raise RuntimeError("topology capacity exceeded")
    path_count = int(step.out.path_count.sum().item())
    ms_plain, _ = timed(step.launch, max(5, args.canyon_steps // 2), 2, stream)
    step.capture()
    ms_batched, _ = timed(step.step, args.canyon_steps, 3, stream)
    value = cells / (ms_batched * 1e-3)

    # each kernel of the step timed alone (HIP events on its stream); the
    # roofline is the longest one's
    fwd_ms = kernel_time(lambda: engine.radiomap_forward(dscene, inputs, topology_capacity=step.capacity,
                                                         out=step.out), stream)
    run_pass = lambda k: (lambda: engine.radiomap_pass(dscene, inputs, step.out, k))  # noqa: E731

    def passes_before(k):
        return lambda: [engine.radiomap_pass(dscene, inputs, step.out, j) for j in range(1, k)]

    chain_ms = kernel_time(run_pass(1), stream)
    main_ms = kernel_time(run_pass(2), stream, before=passes_before(2))
    edge_ms = kernel_time(run_pass(3), stream, before=passes_before(3))
    diff_ms = kernel_time(run_pass(4), stream, before=passes_before(4))
    vjp0_ms = kernel_time(lambda: step.adjoint(1), stream)
    vjp1_ms = kernel_time(lambda: step.adjoint(2), stream, before=lambda: step.adjoint(1))
    reduce_ms = kernel_time(step.reduce, stream)
    n_main, n_diff = descriptor_split(step.out)
    # explicit export of the same solve (counts, device pair offsets, emit,
    # in-pair canonical order, identity compaction, gather of every field),
    # wall time including its two total reads
    engine.export_rows(dscene, inputs, width=depth)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        rows = engine.export_rows(dscene, inputs, width=depth)
    torch.cuda.synchronize()
    export_ms = (time.perf_counter() - t0) * 1e3 / 3
    export_paths = int(rows["tx_id"].shape[0])
    del rows
    public = public_solve_section(args.grid, frequency, depth)
    edges = int(dscene.view.edge_count)
    chains = int(inputs.chains.count)
    scene_bytes = compiled.face_count * 48 + int(dscene.node_count) * 64
    mask_words = lambda n: (n + 31) // 32  # noqa: E731
    kernels = {
        # reads rx + order, writes one chain-validity bit per (cell, chain); chain tables
        "chain_visibility_kernel": (chain_ms, cells * (16 + 4 * mask_words(int(inputs.view.max_tx_chains)))
                                    + scene_bytes + chains * (4 * 5 + 8 * 15)),
        "radiomap_main_tiles": (main_ms, cells * BYTES_CELL_MAIN + n_main * BYTES_PER_DESCRIPTOR + scene_bytes),
        # reads rx + order, writes one visibility bit per (cell, edge); edge table
        "diff_visibility_kernel": (edge_ms, cells * (16 + 4 * mask_words(edges)) + scene_bytes + edges * 80),
        "radiomap_diff_tiles": (diff_ms, cells * BYTES_CELL_DIFF + n_diff * BYTES_PER_DESCRIPTOR + edges * 80),
        "vjp_kernel<0>": (vjp0_ms, cells * BYTES_CELL_VJP + n_main * BYTES_PER_DESCRIPTOR + scene_bytes
                          + step.partials.numel() * 4),
        "vjp_kernel<1>": (vjp1_ms, cells * BYTES_CELL_VJP + n_diff * BYTES_PER_DESCRIPTOR + edges * 80
                          + step.partials.numel() * 8),
    }
    dominant = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dominant]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9

    batch = None
    if not args.no_batch:
        tx_pos, sizes, rx_pos = workloads.batched_transmitters(seed=3, tx_count=args.tx_batch)
        rx_pol = np.tile(np.array([[0, 0, 1]], np.float32), (rx_pos.shape[0], 1))
        tx_pol = np.tile(np.array([[0, 0, 1]], np.float32), (tx_pos.shape[0], 1))
        power = np.ones(tx_pos.shape[0], np.float32)
        binputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                      frequency_hz=frequency, tx_pos=tx_pos, tx_pol=tx_pol, tx_power=power,
                                      rx_pos=rx_pos, rx_pol=rx_pol, rx_count=sizes)
        bstep = MeanDbStep(dscene, binputs, m)
        bstep.capture()
        ms_b, _ = timed(bstep.step, max(5, args.canyon_steps // 2), 2, stream)
        offsets = np.concatenate([[0], np.cumsum(sizes)])
        singles = []
        for t in range(tx_pos.shape[0]):
            sl = slice(int(offsets[t]), int(offsets[t + 1]))
            inp = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                      frequency_hz=frequency, tx_pos=tx_pos[t:t + 1], tx_pol=tx_pol[t:t + 1],
                                      tx_power=power[t:t + 1], rx_pos=rx_pos[sl], rx_pol=rx_pol[sl])
            singles.append(MeanDbStep(dscene, inp, m))

        def plain_loop():
            for s in singles:
                s.launch()

        ms_p, _ = timed(plain_loop, 3, 1, stream)
        bcells = int(sizes.sum())
        batch = {"workload": f"{tx_pos.shape[0]} Tx over the streets (z U[3,30] m), uneven point sets "
                             f"log-uniform 16..4096 (total {bcells}), same scene and physics",
                 "cells": bcells, "batched_ms_per_step": ms_b, "plain_loop_ms_per_step": ms_p,
                 "batched_cells_per_s": bcells / (ms_b * 1e-3), "plain_cells_per_s": bcells / (ms_p * 1e-3),
                 "batched_over_plain": ms_p / ms_b}

    terrain = None
    if not args.no_heightfield:
        # differentiable terrain radiomap: fBm heightfield + boxes, gradient w.r.t. materials, frequency,
        # transmitter and every mesh vertex
        h_freq = 3.5e9
        h_bits = component_bits({"los", "reflection"})
        h_compiled = compile_scene(workloads.heightfield_scene(seed=4), reference_frequency_hz=h_freq)
        h_dscene = device_scene(h_compiled, device)
        h_inputs = engine.solve_inputs(h_compiled, h_dscene, components=h_bits, max_depth=2, coherent=True,
                                       frequency_hz=h_freq)
        h_step = MeanDbStep(h_dscene, h_inputs, h_compiled.materials.count, vertices=True)
        if bool((h_step.out.failure != 0).any()):
            raise RuntimeError("topology capacity exceeded (heightfield)")
        h_step.capture()
        h_ms, _ = timed(h_step.step, max(5, args.canyon_steps // 2), 2, stream)
        h_geo_ms = kernel_time(h_step.geometry, stream)
        h_fwd_ms = kernel_time(lambda: engine.radiomap_forward(h_dscene, h_inputs, topology_capacity=h_step.capacity,
                                                               out=h_step.out), stream)
        h_cells = h_inputs.total_cells
        # geometry step of a vertex loop on the same mesh (every vertex lifted by up to 5 cm): device refit
        # (triangle tables, tree boxes, planes, edges; wall time incl. the failure-word read) against a host
        # recompile of the moved scene plus a fresh SAH tree and upload
        lift = np.random.default_rng(1).uniform(0.0, 0.05, h_compiled.vertices.shape[0]).astype(np.float32)
        moved = h_compiled.vertices.copy()
        moved[:, 2] += lift
        refit_scene = DeviceScene(h_compiled, device)
        moved_t = torch.as_tensor(moved).to(device)
        refit_scene.refit(moved_t)
        torch.cuda.synchronize()
        r0 = time.perf_counter()
        for _ in range(5):
            refit_scene.refit(moved_t)
        torch.cuda.synchronize()
        refit_ms = (time.perf_counter() - r0) * 1e3 / 5
        r0 = time.perf_counter()
        for _ in range(2):
            DeviceScene(with_vertices(h_compiled, moved), device)
        torch.cuda.synchronize()
        rebuild_ms = (time.perf_counter() - r0) * 1e3 / 2
        del refit_scene
        terrain = {"workload": f"fBm heightfield 128x128 at 2 m (relief 20 m) + 40 boxes, {h_compiled.face_count} "
                               "tris, 4 materials, 256x256 grid 1.5 m above the highest point, 3.5 GHz, LoS + "
                               "depth-2 reflection, loss mean dB; VJP w.r.t. eps_r, sigma, f, Tx and all "
                               f"{h_compiled.vertices.shape[0]} vertices",
                   "cells": h_cells, "chains": h_inputs.chains.count,
                   "paths": int(h_step.out.path_count.sum().item()), "ms_per_step": h_ms,
                   "cells_per_s": h_cells / (h_ms * 1e-3), "forward_ms": h_fwd_ms, "vertex_vjp_ms": h_geo_ms,
                   "geometry_update": {"device_refit_ms": refit_ms, "host_recompile_and_build_ms": rebuild_ms}}

    paths_batch = None if args.no_batch else batched_paths_section(stream)
    canyon_cpu = None
    if not args.no_cpu:
        from tests.gpu_helpers import Run, run_oracle

        rng = np.random.default_rng(0)
        sample = np.sort(rng.choice(cells, min(cells, 16384), replace=False))
        run = Run(compiled, dscene, inputs, step.out, bits, depth, frequency)
        c0 = time.perf_counter()
        ref = run_oracle(run, 0, sample)
        cpu_s = time.perf_counter() - c0
        p_gpu = step.out.power_total.cpu().numpy()[sample]
        p_ref = ref["record"][:, 0]
        both = (p_gpu > 0) & (p_ref > 0)
        err = np.abs(10 * np.log10(p_gpu[both] / p_ref[both]))
        canyon_cpu = {"cells_per_s": len(sample) / cpu_s, "sample": f"{len(sample)} random cells, {cpu_s:.1f} s",
                      "cells_lit_one_only": int(((p_gpu > 0) != (p_ref > 0)).sum()),
                      "median_db_error_vs_gpu": float(np.median(err)) if err.size else 0.0,
                      "max_db_error_vs_gpu": float(err.max()) if err.size else 0.0,
                      "cells_above_0.01_db": int((err > 0.01).sum())}

    canyon = {
        "cells": cells, "paths": path_count, "chains": inputs.chains.count,
        "ms_per_step": ms_batched, "cells_per_s": cells / (ms_batched * 1e-3),
        "plain_ms_per_step": ms_plain, "forward_ms": fwd_ms,
        "kernel_ms": {**{k: v[0] for k, v in kernels.items()}, "vjp_reduce_kernel": reduce_ms},
        "paths_reflection_pass": n_main, "paths_diffraction": n_diff, "chain_planning_s": plan_s,
        "export": {"ms": export_ms, "paths": export_paths, "paths_per_s": export_paths / (export_ms * 1e-3)},
        "public_solve": public,
        "dominant_kernel": dominant, "dominant_hbm_frac": (dom_bytes / (dom_ms * 1e-3) / 1e9) / HBM_PEAK_GBS,
        "oracle_check": canyon_cpu,
    }
    if city is None:
        print(json.dumps({"canyon_only": True, **canyon, "multi_tx": batch, "batched_paths": paths_batch,
                          "heightfield": terrain}))
        return 0
    canyon_name = (f"urban canyon 8x8 boxes, {compiled.face_count} tris, 1 Tx at street intersection "
                   f"(53.231,52.013,25), {args.grid}x{args.grid} grid at 2 m, 28 GHz, depth 3 refl + UTD diff + "
                   "LoS, loss mean dB over lit cells")
    workload_name = (f"scaled city seed 5: 400 boxes + 256x256 fBm heightfield ({city['triangles']} tris), 8 materials, "
                     f"{city['tx']} Tx with grids 64^2..512^2 ({city['grid_cells']} cells), 5-bounce reflection + "
                     "UTD diffraction + transmission + LoS at 3.5 and 28 GHz, loss mean dB over lit cells, VJP w.r.t. "
                     "materials, frequencies and Tx positions")
    value = city["cells"] / (city["ms_batched"] * 1e-3)
    achieved = city["dom_bytes"] / (city["dom_ms"] * 1e-3) / 1e9
    line = {
        "metric": "radiomap cells/s fwd+VJP (batched vs plain, 1 GPU) at <=0.01 dB err vs FP64",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": city["ms_batched"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry/fields/tangents, f64 phase, UTD transition and cell sums",
        "data": "synthetic seeded scaled city (seed 5)",
        "config": {
            "workload": workload_name,
            "pass": "forward+vjp",
            "cells": city["cells"],
            "cells_per_carrier": city["grid_cells"],
            "carriers_hz": city["frequencies"],
            "paths": city["paths"],
            "paths_per_s": city["paths"] / (city["ms_batched"] * 1e-3),
            "chains": city["chains"],
            "edges": city["edges"],
            "plain_ms_per_step": city["ms_plain"],
            "plain_cells_per_s": city["cells"] / (city["ms_plain"] * 1e-3) if city["ms_plain"] else None,
            "batched_over_plain": city["ms_plain"] / city["ms_batched"] if city["ms_plain"] else None,
            "batched_maps_bit_identical_to_plain": city["identical"],
            "launch_shape": city["shape"],
            "batches": city["batches"],
            "autotune_s": city["tune_s"],
            "chain_planning_s": city["planning_s"],
            "workspace_gb": city["workspace_gb"],
            "gradients": city["grads"],
            "heaviest_batch": {"tx": city["unit_tx"], "cells": city["unit_cells"], "kernel_ms": city["unit_ms"]},
            "canyon": {"workload": canyon_name, **canyon},
            "multi_tx": batch,
            "batched_paths": paths_batch,
            "heightfield": terrain,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_bytes("chain_validity_pass", workload_name),
                     "kernel": "chain_visibility_seg_kernel (with its pass's super_box, super_cull and fill_zero "
                               "launches; average over the batches)",
                     "kernel_ms": city["dom_ms"], "algorithmic_bytes_per_launch": city["dom_bytes"]},
        "cpu_baseline": city["cpu"],
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
