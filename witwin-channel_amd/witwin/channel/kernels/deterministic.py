"""Differentiable radiomap: ``torch.autograd.Function`` over the fused kernels.

Forward runs ``wc_radiomap_forward`` with path descriptors stored per cell;
backward runs ``wc_radiomap_vjp`` (the fixed-topology adjoint that recomputes
each stored path with dual numbers) and reduces its per-tile partials.
Differentiable inputs: first-layer ``eps_r`` and ``sigma_e`` of every
material, the carrier frequency, and the transmitter positions.

Counterpart of the reference ``_DeterministicAccumulateFlatAdFunction`` and
the field AD functions (witwin/channel/kernels/deterministic.py:186-390,
witwin/channel/kernels/fields.py:1087-1827), which keep per-path tapes; here
nothing is taped beyond the per-cell descriptors (4 bytes per path).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

import torch

from witwin.channel import _native
from witwin.channel.propagation import engine

DEFAULT_TOPOLOGY_CAPACITY = 64


@dataclass
class RadiomapState:
    dscene: Any
    inputs: engine.SolveInputs
    material_count: int
    topology_capacity: int = DEFAULT_TOPOLOGY_CAPACITY
    out: engine.RadiomapOutputs | None = None
    partials: torch.Tensor | None = None


class RadiomapFunction(torch.autograd.Function):
    """outputs: power_total [cells], field_total [cells] c64,
    component_power [4, cells], component_field [4, cells] c64,
    diffraction_power [cells] (= sum over axes of |diffraction vector|^2)."""

    @staticmethod
    def forward(ctx, eps_r, sigma_e, frequency, tx_position, state: RadiomapState):  # type: ignore[override]
        state.out = engine.radiomap_forward(state.dscene, state.inputs, topology_capacity=state.topology_capacity,
                                            out=state.out)
        out = state.out
        diff_power = out.diffraction_vector.abs().square().sum(dim=-1)
        ctx.state = state
        return out.power_total, out.field_total, out.component_power, out.component_field, diff_power

    @staticmethod
    def backward(ctx, g_power, g_field, g_cpower, g_cfield, g_dpower):  # type: ignore[override]
        state: RadiomapState = ctx.state
        inputs = state.inputs
        out = state.out
        device = out.power_total.device
        m = state.material_count
        stride = int(_native.call("wc_vjp_partial_stride", m))
        tiles = inputs.tile_count
        if state.partials is None or state.partials.shape != (tiles, stride):
            state.partials = torch.empty((tiles, stride), dtype=torch.float32, device=device)

        def prep(g, dtype):
            if g is None:
                return None
            return g.to(dtype).contiguous()

        gp = prep(g_power, torch.float32)
        gf = prep(g_field, torch.complex64)
        gcp = prep(g_cpower, torch.float32)
        gcf = prep(g_cfield, torch.complex64)
        gd = prep(g_dpower, torch.float32)
        p = _native.ptr
        view = _native.WcAdjoint(
            grad_power_total=p(gp), grad_field_total=p(gf), grad_component_power=p(gcp),
            grad_component_field=p(gcf), grad_diffraction_power=p(gd), power_total=p(out.power_total),
            field_total=p(out.field_total), component_field=p(out.component_field),
            diffraction_vector=p(out.diffraction_vector), topology=p(out.topology), path_count=p(out.path_count),
            topology_capacity=out.topology_capacity, partials=p(state.partials),
        )
        _native.call("wc_radiomap_vjp", state.dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(view),
                     p(inputs.tiles), tiles, _native.stream_handle(device))
        part = state.partials
        grad_eps = part[:, :m].sum(dim=0)
        grad_sigma = part[:, m:2 * m].sum(dim=0)
        grad_freq = part[:, 2 * m].sum()
        tile_tx = inputs.tiles[:tiles, 0].long()
        grad_tx = torch.zeros((int(inputs.view.tx_count), 3), dtype=torch.float32, device=device)
        grad_tx.index_add_(0, tile_tx, part[:, 2 * m + 1:2 * m + 4])
        return grad_eps, grad_sigma, grad_freq, grad_tx, None


def radiomap_vjp_outputs(dscene, inputs: engine.SolveInputs, *, eps_r: torch.Tensor, sigma_e: torch.Tensor,
                         frequency: torch.Tensor, tx_position: torch.Tensor, material_count: int,
                         state: RadiomapState | None = None, topology_capacity: int = DEFAULT_TOPOLOGY_CAPACITY):
    """Differentiable radiomap outputs (see RadiomapFunction) for explicit
    parameter tensors. The tensors' values must equal the compiled scene's."""
    if state is None:
        state = RadiomapState(dscene, inputs, material_count, topology_capacity)
    outs = RadiomapFunction.apply(eps_r, sigma_e, frequency, tx_position, state)
    return outs, state


def _leaf_vector(values: list, fallback: torch.Tensor, device) -> torch.Tensor:
    items = []
    for i, v in enumerate(values):
        if isinstance(v, torch.Tensor):
            items.append(v.reshape(()).to(device=device, dtype=torch.float32))
        else:
            items.append(fallback[i].to(device))
    return torch.stack(items) if items else fallback.to(device)


def radiomap_with_vjp(compiled, dscene, inputs, *, material_params: Any = None, tx_position: Any = None):
    """Used by deterministic.solve(ad_mode="vjp"): gathers the AD leaves the
    compile recorded (material eps_r / sigma_e tensors, transmitter position
    tensors, a tensor frequency) and runs the differentiable radiomap."""
    device = dscene.device
    leaves = compiled.cache.get("ad_leaves", {})
    m = compiled.materials
    eps = _leaf_vector(leaves.get("eps_r", []), torch.as_tensor(m.eps_r), device)
    sig = _leaf_vector(leaves.get("sigma_e", []), torch.as_tensor(m.sigma_e), device)
    freq_leaf = leaves.get("frequency")
    freq = freq_leaf.reshape(()).to(device=device, dtype=torch.float32) if isinstance(freq_leaf, torch.Tensor) \
        else torch.tensor(compiled.frequency_hz, dtype=torch.float32, device=device)
    tx_leaves = leaves.get("tx_position", [])
    if tx_position is not None:
        txp = tx_position
    elif tx_leaves:
        txp = torch.stack([t.reshape(3).to(device=device, dtype=torch.float32) for t in tx_leaves])
    else:
        txp = inputs.tx_pos
    outs, state = radiomap_vjp_outputs(dscene, inputs, eps_r=eps, sigma_e=sig, frequency=freq, tx_position=txp,
                                       material_count=m.count)
    power, field, comp_power, comp_field, diff_power = outs
    return state.out, (power, field, comp_power, diff_power)
