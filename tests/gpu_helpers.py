"""Shared helpers of the GPU parity tests: run a scene through the HIP engine
and through the FP64 oracle on the same inputs."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

import oracle
from witwin.channel.components import component_bits
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import compile as compile_scene, device_scene

# |10 log10(P_gpu / P_oracle)| gate on path / cell power (dB) and coefficient phase (rad)
DB_TOL = 0.01
PHASE_TOL = 1.0e-3


@dataclass
class Run:
    compiled: object
    dscene: object
    inputs: object
    out: object
    bits: int
    depth: int
    frequency: float


def run_gpu(scene, components, max_depth, frequency, device, *, coherent=True, topology_capacity=0,
            tiled=True, ray_count=engine.DISCOVERY_RAY_COUNT) -> Run:
    compiled = compile_scene(scene, reference_frequency_hz=frequency)
    dscene = device_scene(compiled, device)
    bits = component_bits(components)
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=max_depth, coherent=coherent,
                                 frequency_hz=frequency, ray_count=ray_count)
    out = engine.radiomap_forward(dscene, inputs, topology_capacity=topology_capacity, tiled=tiled)
    torch.cuda.synchronize()
    return Run(compiled, dscene, inputs, out, bits, max_depth, frequency)


def tx_chains(run: Run, t: int) -> tuple[np.ndarray, np.ndarray]:
    off = run.inputs.chains.offset_host
    lo, hi = off[t * 5], off[t * 5 + 5]
    groups = run.inputs.chains.groups_host[lo:hi]
    depth = (groups >= 0).sum(axis=1).astype(np.int32)
    return groups, depth


def run_oracle(run: Run, t: int, rx_index: np.ndarray, *, coherent=True, paths_cap=0, perturb=0.0):
    ep = run.compiled.endpoints
    groups, depth = tx_chains(run, t)
    orc = oracle.Oracle(run.compiled)
    return orc.radiomap(tx=ep.tx_position[t], tx_pol=ep.tx_polarization[t], tx_power=float(ep.tx_power[t]),
                        rx=ep.rx_position[rx_index], rx_pol=ep.rx_polarization[rx_index[0]] if len(rx_index) else
                        (0, 0, 1), frequency=run.frequency, components=run.bits, max_depth=run.depth,
                        chains=groups, chain_depth=depth, coherent=coherent, paths_cap=paths_cap, perturb=perturb)


def db_error(p_gpu: np.ndarray, p_ref: np.ndarray) -> np.ndarray:
    p_gpu = np.asarray(p_gpu, dtype=np.float64)
    p_ref = np.asarray(p_ref, dtype=np.float64)
    both = (p_gpu > 0) & (p_ref > 0)
    err = np.zeros_like(p_ref)
    err[both] = np.abs(10.0 * np.log10(p_gpu[both] / p_ref[both]))
    err[(p_gpu > 0) != (p_ref > 0)] = np.inf
    return err


def phase_error(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.abs(np.angle(np.asarray(a, np.complex128) * np.conj(np.asarray(b, np.complex128))))


def coherence_condition(record: np.ndarray, paths) -> np.ndarray:
    """|sum F|^2 / sum |F|^2 per cell: small values mark deep interference nulls,
    where the relative error of the coherent power is amplified by the inverse
    of this ratio."""
    total = record[:, 0]
    incoherent = np.array([sum(p.gain for p in cell) for cell in paths]) if paths is not None else None
    return total / np.maximum(incoherent, 1e-300)
