"""Radiomap forward+VJP throughput on one MI355X.

Workload (the value line): seeded urban canyon (8 x 8 box buildings, 770
triangles, 3 materials incl. ground), one transmitter at a street
intersection (53.231, 52.013, 25) m,
28 GHz, up to 3 specular bounces + first-order UTD diffraction + line of sight,
256 x 256 receiver grid at 2 m, z = 1.5 m. One step = radiomap forward (fused
discovery + fields + float64 coherent cell sums, path descriptors stored) +
the fixed-topology VJP of L = mean dB path gain over lit cells w.r.t. every
material's eps_r and sigma, the frequency and the transmitter position, with
inputs resident in HBM. ``value`` = cells / step time of the batched form
(persistent wave-tiled launches captured once in a HIP graph and replayed).

Also reported: the plain form (the same kernels launched eagerly per step),
a 64-transmitter batched solve with uneven receiver sets (log-uniform 16..4096
street points per transmitter) against a per-transmitter loop, the roofline of
the dominant kernel, and the FP64 CPU oracle on the same cells.

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
BYTES_PER_CELL_FWD = 48.0      # path_gain f32 + field c64 + 3 x (component power f32 + field c64)
BYTES_PER_CELL_VJP = 4.0       # dL/dP_cell read by the adjoint
BYTES_PER_TRIANGLE = 48.0      # scene bytes read once per launch


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


def kernel_time(fn, stream, reps: int = 5) -> float:
    """Average duration (ms) of one launch sequence bracketed by HIP events."""
    import torch

    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--grid", type=int, default=256)
    parser.add_argument("--tx-batch", type=int, default=64)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--no-batch", action="store_true")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(the path shards by transmitter, see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from witwin.channel import workloads
    from witwin.channel.components import component_bits
    from witwin.channel.kernels.deterministic import MeanDbStep
    from witwin.channel.propagation import engine
    from witwin.channel.scene.compiler import compile as compile_scene, device_scene

    device = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(device)
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
    ms_plain, _ = timed(step.launch, max(5, args.steps // 2), 2, stream)
    step.capture()
    ms_batched, _ = timed(step.step, args.steps, args.warmup, stream)
    value = cells / (ms_batched * 1e-3)

    # dominant kernel: the fused forward (fwd) and the adjoint, each timed alone
    fwd_ms = kernel_time(lambda: engine.radiomap_forward(dscene, inputs, topology_capacity=step.capacity,
                                                         out=step.out), stream)
    traffic_fwd = cells * BYTES_PER_CELL_FWD + compiled.face_count * BYTES_PER_TRIANGLE
    achieved = traffic_fwd / (fwd_ms * 1e-3) / 1e9

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
        ms_b, _ = timed(bstep.step, max(5, args.steps // 2), 2, stream)
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

    cpu = None
    if not args.no_cpu:
        from tests.gpu_helpers import Run, run_oracle

        rng = np.random.default_rng(0)
        sample = np.sort(rng.choice(cells, min(cells, 16384), replace=False))
        run = Run(compiled, dscene, inputs, step.out, bits, depth, frequency)
        c0 = time.perf_counter()
        ref = run_oracle(run, 0, sample)
        cpu_s = time.perf_counter() - c0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        p_gpu = step.out.power_total.cpu().numpy()[sample]
        p_ref = ref["record"][:, 0]
        both = (p_gpu > 0) & (p_ref > 0)
        err = np.abs(10 * np.log10(p_gpu[both] / p_ref[both]))
        cpu = {"value": len(sample) / cpu_s, "unit": "cells/s (forward only)", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} random cells of the same 256x256 radiomap; FP64 oracle with brute-force "
                         f"visibility, OpenMP over cells, {cpu_s:.1f} s",
               "median_db_error_vs_gpu": float(np.median(err)) if err.size else 0.0,
               "p99_db_error_vs_gpu": float(np.percentile(err, 99)) if err.size else 0.0}

    line = {
        "metric": "radiomap cells/s fwd+VJP (batched vs plain, 1 GPU) at <=0.01 dB err vs FP64",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batched,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry/fields/tangents, f64 phase, UTD transition and cell sums",
        "data": "synthetic seeded urban canyon (seed 2)",
        "config": {
            "workload": f"urban canyon 8x8 boxes, {compiled.face_count} tris, 1 Tx at street intersection (53.231,52.013,25), "
                        f"{args.grid}x{args.grid} grid at 2 m, 28 GHz, depth 3 refl + UTD diff + LoS, "
                        "loss mean dB over lit cells",
            "pass": "forward+vjp",
            "cells": cells,
            "paths": path_count,
            "chains": inputs.chains.count,
            "plain_ms_per_step": ms_plain,
            "plain_cells_per_s": cells / (ms_plain * 1e-3),
            "batched_over_plain": ms_plain / ms_batched,
            "forward_kernel_ms": fwd_ms,
            "chain_planning_s": plan_s,
            "multi_tx": batch,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "radiomap_tiles_kernel", "kernel_ms": fwd_ms,
                     "algorithmic_bytes_per_launch": traffic_fwd},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
