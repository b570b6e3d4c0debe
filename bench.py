"""Radiomap throughput benchmark on one MI355X.

Workload: seeded urban canyon (8 x 8 box buildings, ~780 triangles, 2
materials), one transmitter at (22.5, 22.5, 25) m, 28 GHz, up to 3 specular
bounces + first-order UTD diffraction + line of sight, 256 x 256 receiver grid
at 2 m, z = 1.5 m. One step = one radiomap forward (+ the fixed-topology VJP of
L = mean dB path gain when available) over all 65,536 cells with inputs
resident in HBM. Reported: cells/s of the batched (persistent, wave-tiled)
launch, the plain per-transmitter launch for comparison, the roofline of the
dominant kernel, and the FP64 CPU oracle on a bounded sample of cells.

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


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--grid", type=int, default=256)
    parser.add_argument("--cpu-sample", type=int, default=48)
    parser.add_argument("--no-cpu", action="store_true")
    args = parser.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(the path shards by transmitter, see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from witwin.channel import workloads
    from witwin.channel.components import component_bits
    from witwin.channel.propagation import engine
    from witwin.channel.scene.compiler import compile as compile_scene, device_scene

    device = torch.device("cuda", 0)
    frequency = 28.0e9
    depth = 3
    components = {"los", "reflection", "diffraction"}
    scene = workloads.urban_canyon_scene(seed=2, shape=(args.grid, args.grid), spacing=2.0)
    compiled = compile_scene(scene, reference_frequency_hz=frequency)
    dscene = device_scene(compiled, device)
    bits = component_bits(components)
    t0 = time.perf_counter()
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                 frequency_hz=frequency)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    cells = inputs.total_cells
    out = engine.radiomap_forward(dscene, inputs)
    stream = torch.cuda.current_stream(device)

    def step(tiled: bool) -> None:
        engine.radiomap_forward(dscene, inputs, tiled=tiled, out=out)

    def timed(tiled: bool, steps: int, warmup: int) -> tuple[float, float]:
        for _ in range(warmup):
            step(tiled)
        torch.cuda.synchronize()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        wall0 = time.perf_counter()
        for i in range(steps):
            starts[i].record(stream)
            step(tiled)
            ends[i].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
        return wall / steps * 1e3, kernel_ms

    ms_batched, kernel_ms = timed(True, args.steps, args.warmup)
    ms_plain, _ = timed(False, max(3, args.steps // 4), 1)
    value = cells / (ms_batched * 1e-3)
    traffic_bytes = cells * BYTES_PER_CELL_FWD + compiled.face_count * BYTES_PER_TRIANGLE
    achieved = traffic_bytes / (kernel_ms * 1e-3) / 1e9
    path_count = int(out.path_count.sum().item())

    cpu = None
    if not args.no_cpu:
        import oracle

        from tests.gpu_helpers import Run, run_oracle  # noqa: F401 - helper shared with the parity tests

        rng = np.random.default_rng(0)
        sample = np.sort(rng.choice(cells, min(args.cpu_sample, cells), replace=False))
        run = Run(compiled, dscene, inputs, out, bits, depth, frequency)
        c0 = time.perf_counter()
        ref = run_oracle(run, 0, sample)
        cpu_s = time.perf_counter() - c0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": len(sample) / cpu_s, "unit": "cells/s", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} random cells of the same 256x256 radiomap, FP64 oracle, brute-force "
                         f"visibility, OpenMP ({cpu_s:.1f} s)"}
        p_gpu = out.power_total.cpu().numpy()[sample]
        with np.errstate(divide="ignore", invalid="ignore"):
            err = np.abs(10 * np.log10(p_gpu / ref["record"][:, 0]))
        err = err[np.isfinite(err)]
        cpu["max_db_error_vs_gpu"] = float(err.max()) if err.size else 0.0

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
        "dtype": "f32 geometry/fields, f64 phase and cell sums",
        "data": "synthetic seeded urban canyon (seed 2)",
        "config": {
            "workload": f"urban canyon 8x8 boxes, {compiled.face_count} tris, 1 Tx, {args.grid}x{args.grid} grid "
                        "at 2 m, 28 GHz, depth 3 refl + UTD diff + LoS",
            "pass": "forward",
            "cells": cells,
            "paths": path_count,
            "chains": inputs.chains.count,
            "plain_ms_per_step": ms_plain,
            "plain_cells_per_s": cells / (ms_plain * 1e-3),
            "batched_over_plain": ms_plain / ms_batched,
            "chain_planning_s": plan_s,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "radiomap_tiles_kernel", "kernel_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": traffic_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
