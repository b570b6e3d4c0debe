"""Differentiable radiomap: ``torch.autograd.Function`` over the fused kernels.

Forward runs ``wc_radiomap_forward`` with path descriptors stored per cell;
backward runs ``wc_radiomap_vjp`` (the fixed-topology adjoint that recomputes
each stored path with dual numbers) and reduces its per-tile partials; the
forward-mode ``jvp`` runs ``wc_radiomap_jvp`` (the same recomputation with one
tangent seeded by the input tangents).
Differentiable inputs: first-layer ``eps_r`` and ``sigma_e`` of every
material, the carrier frequency, the transmitter positions and the mesh
vertices (``wc_radiomap_vjp_geometry`` + ``wc_geometry_vertex_grad``, run
only when the vertex table requires a gradient).

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
from witwin.channel.kernels._accumulate import (  # noqa: F401  (reference facade names)
    deterministic_accumulate_flat, deterministic_accumulate_flat_ad, deterministic_accumulate_flat_backward,
    deterministic_accumulate_flat_fwd64, deterministic_accumulate_flat_jvp)
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
    def forward(ctx, eps_r, sigma_e, frequency, tx_position, vertices, state: RadiomapState):  # type: ignore[override]
        apply_parameters(state, eps_r, sigma_e, frequency, tx_position, vertices)
        if state.out is None:
            state.out = engine.forward_with_topology(state.dscene, state.inputs, capacity=state.topology_capacity)
            state.topology_capacity = state.out.topology_capacity
        else:
            state.out = engine.radiomap_forward(state.dscene, state.inputs,
                                                topology_capacity=state.topology_capacity, out=state.out)
        out = state.out
        diff_power = out.diffraction_vector.abs().square().sum(dim=-1)
        ctx.state = state
        return out.power_total, out.field_total, out.component_power, out.component_field, diff_power

    @staticmethod
    def jvp(ctx, d_eps, d_sigma, d_frequency, d_tx, d_vertices, _d_state):  # type: ignore[override]
        """Forward-mode tangents of the five outputs (wc_radiomap_jvp): the
        stored paths recomputed with one-tangent dual numbers seeded by the
        input tangents (first-layer eps_r / sigma_e, frequency, transmitter
        positions). Vertex tangents are not available in forward mode."""
        if d_vertices is not None and bool((d_vertices != 0).any()):
            raise NotImplementedError("forward-mode tangents of the mesh vertices are not available; use "
                                      "ad_mode='vjp' for vertex gradients")
        return radiomap_jvp(ctx.state, d_eps, d_sigma, d_frequency, d_tx)

    @staticmethod
    def backward(ctx, g_power, g_field, g_cpower, g_cfield, g_dpower):  # type: ignore[override]
        state: RadiomapState = ctx.state
        inputs = state.inputs
        out = state.out
        device = out.power_total.device
        m = state.material_count
        stride = int(_native.call("wc_vjp_partial_stride", m))
        tiles = inputs.tile_count
        rows = int(_native.call("wc_vjp_partial_rows", tiles))
        if state.partials is None or state.partials.shape != (rows, stride):
            state.partials = torch.empty((rows, stride), dtype=torch.float32, device=device)

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
        stream = _native.stream_handle(device)
        _native.call("wc_radiomap_vjp", state.dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(view),
                     p(inputs.tiles), tiles, stream)
        t = int(inputs.view.tx_count)
        grad_params = torch.empty(2 * m + 1, dtype=torch.float32, device=device)
        grad_tx = torch.empty((t, 3), dtype=torch.float32, device=device)
        _native.call("wc_vjp_reduce", p(state.partials), tiles, stride, p(inputs.tiles), t, p(grad_params),
                     p(grad_tx), stream)
        grad_vertices = geometry_vjp(state.dscene, inputs, view) if ctx.needs_input_grad[4] else None
        return grad_params[:m], grad_params[m:2 * m], grad_params[2 * m], grad_tx, grad_vertices, None


def radiomap_jvp(state: RadiomapState, d_eps=None, d_sigma=None, d_frequency=None, d_tx=None):
    """Tangents of (power_total, field_total, component_power,
    component_field, diffraction_power) of the last forward of ``state``
    along the given input tangents (None = zero)."""
    out = state.out
    inputs = state.inputs
    device = out.power_total.device
    m = state.material_count
    cells = inputs.total_cells

    def vec(x, n):
        if x is None:
            return None
        return x.detach().to(device=device, dtype=torch.float32).reshape(-1)[:n].contiguous()

    de, ds = vec(d_eps, m), vec(d_sigma, m)
    dt = vec(d_tx, 3 * int(inputs.view.tx_count))
    df = 0.0 if d_frequency is None else float(d_frequency.detach().reshape(()))
    p = _native.ptr
    d_power = torch.empty(cells, dtype=torch.float32, device=device)
    d_field = torch.empty(cells, dtype=torch.complex64, device=device)
    d_cpower = torch.empty((4, cells), dtype=torch.float32, device=device)
    d_cfield = torch.empty((4, cells), dtype=torch.complex64, device=device)
    d_dpower = torch.empty(cells, dtype=torch.float32, device=device)
    primal = _native.WcAdjoint(
        grad_power_total=None, grad_field_total=None, grad_component_power=None, grad_component_field=None,
        grad_diffraction_power=None, power_total=p(out.power_total), field_total=p(out.field_total),
        component_field=p(out.component_field), diffraction_vector=p(out.diffraction_vector),
        topology=p(out.topology), path_count=p(out.path_count), topology_capacity=out.topology_capacity,
        partials=None)
    tin = _native.WcTangentIn(d_eps_r=p(de), d_sigma_e=p(ds), d_frequency=df, d_tx_pos=p(dt))
    tout = _native.WcTangentOut(d_power_total=p(d_power), d_field_total=p(d_field), d_component_power=p(d_cpower),
                                d_component_field=p(d_cfield), d_diffraction_power=p(d_dpower))
    _native.call("wc_radiomap_jvp", state.dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(primal),
                 ctypes.byref(tin), ctypes.byref(tout), _native.stream_handle(device))
    return d_power, d_field, d_cpower, d_cfield, d_dpower


def apply_parameters(state: RadiomapState, eps_r, sigma_e, frequency, tx_position, vertices) -> None:
    """Write the differentiable inputs' current values into the device scene
    and the solve before a forward: first-layer eps_r / sigma_e of every
    material (scalar view and layer table), the carrier frequency, the
    transmitter positions (the reflection chains keep their group sequences;
    their image points are recomputed) and, when a vertex table other than
    the scene's own is given, the vertices (device refit, same topology)."""
    d, inputs = state.dscene, state.inputs
    m = state.material_count
    first = d.layer_offset[:m].long()
    eps = eps_r.detach().to(device=d.device, dtype=torch.float32).reshape(-1)[:m]
    sig = sigma_e.detach().to(device=d.device, dtype=torch.float32).reshape(-1)[:m]
    d.mat_eps[:m].copy_(eps)
    d.mat_sigma[:m].copy_(sig)
    d.layer_eps.index_copy_(0, first, eps)
    d.layer_sigma.index_copy_(0, first, sig)
    inputs.view.frequency_hz = float(frequency.detach()) if isinstance(frequency, torch.Tensor) else float(frequency)
    txp = tx_position.detach().to(device=d.device, dtype=torch.float32).reshape(inputs.tx_pos.shape)
    moved = False
    if txp.data_ptr() != inputs.tx_pos.data_ptr():
        inputs.tx_pos.copy_(txp)
        moved = True
    if vertices is not None and vertices.data_ptr() != d.vertices.data_ptr():
        d.refit(vertices)
        moved = True
    if moved:
        # images (and receiver cones) follow the transmitters and the group planes
        engine.refresh_chain_geometry(d, inputs)


def geometry_vjp(dscene, inputs: engine.SolveInputs, view) -> torch.Tensor:
    """dL/d(vertices) [V, 3] of the fixed-topology radiomap for the adjoint
    ``view``: per-group plane and per-edge accumulators from
    wc_radiomap_vjp_geometry, then the vertex Jacobians and per-vertex sums."""
    device = dscene.device
    p = _native.ptr
    g, e = dscene.group_count, dscene.edge_count
    if e and int(inputs.view.components) & 4:
        raise RuntimeError("vertex gradients of diffraction paths are not available; solve without diffraction "
                           "(components without 'diffraction') to differentiate the mesh vertices")
    plane = torch.zeros((g, 4), dtype=torch.float64, device=device)
    edge = torch.zeros((e, 8), dtype=torch.float64, device=device)
    stream = _native.stream_handle(device)
    _native.call("wc_radiomap_vjp_geometry", dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(view),
                 p(inputs.tiles), inputs.tile_count, p(plane), p(edge), stream)
    offset, slot = dscene.vertex_csr()
    v = int(dscene.vertices.shape[0])
    contrib = torch.empty(9 * g + 12 * e, dtype=torch.float32, device=device)
    grad = torch.empty((v, 3), dtype=torch.float32, device=device)
    _native.call("wc_geometry_vertex_grad", p(dscene.vertices), p(dscene.faces), p(dscene.group_face), g, p(plane),
                 p(dscene.edge_vertices), e, p(edge), p(offset), p(slot), v, p(contrib), p(grad), stream)
    return grad


def radiomap_vjp_outputs(dscene, inputs: engine.SolveInputs, *, eps_r: torch.Tensor, sigma_e: torch.Tensor,
                         frequency: torch.Tensor, tx_position: torch.Tensor, material_count: int,
                         vertices: torch.Tensor | None = None, state: RadiomapState | None = None,
                         topology_capacity: int = DEFAULT_TOPOLOGY_CAPACITY):
    """Differentiable radiomap outputs (see RadiomapFunction) for explicit
    parameter tensors. The tensors' values must equal the compiled scene's
    (``vertices``: the concatenated [V, 3] vertex table)."""
    if state is None:
        state = RadiomapState(dscene, inputs, material_count, topology_capacity)
    if vertices is None:
        vertices = dscene.vertices
    outs = RadiomapFunction.apply(eps_r, sigma_e, frequency, tx_position, vertices, state)
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
    vert_leaves = leaves.get("vertices", [])
    vertices = None
    if any(isinstance(t, torch.Tensor) and t.requires_grad for t, _ in vert_leaves):
        # the concatenated vertex table, differentiable back to every structure's vertices
        vertices = torch.cat([t.reshape(-1, 3).to(device=device, dtype=torch.float32) for t, _ in vert_leaves])
    outs, state = radiomap_vjp_outputs(dscene, inputs, eps_r=eps, sigma_e=sig, frequency=freq, tx_position=txp,
                                       material_count=m.count, vertices=vertices)
    power, field, comp_power, comp_field, diff_power = outs
    return state.out, (power, field, comp_power, diff_power)


class MeanDbStep:
    """One forward + VJP step of L = mean over lit cells of 10 log10(P_cell),
    written with launches only (no host synchronization, no allocation after
    the first call) so it can be captured once in a HIP graph and replayed.

    dL/dP_c = 10 / (ln 10 * N * P_c) on the N cells lit in the first forward
    (the fixed topology keeps that set). Produces ``grad_params`` =
    [dL/deps_r (M), dL/dsigma (M), dL/df] and ``grad_tx`` [T, 3], and with
    ``vertices=True`` also ``grad_vertices`` [V, 3].
    """

    def __init__(self, dscene, inputs: engine.SolveInputs, material_count: int,
                 topology_capacity: int = DEFAULT_TOPOLOGY_CAPACITY, vertices: bool = False,
                 workspace: torch.Tensor | None = None) -> None:
        self.dscene = dscene
        self.inputs = inputs
        self.m = material_count
        self.capacity = topology_capacity
        device = dscene.device
        self.out = engine.forward_with_topology(dscene, inputs, capacity=topology_capacity, workspace=workspace)
        self.capacity = self.out.topology_capacity
        power = self.out.power_total
        lit = power > 0
        count = max(int(lit.sum().item()), 1)
        self.scale = torch.where(lit, torch.full_like(power, 10.0 / (2.302585092994046 * count)),
                                 torch.zeros_like(power))
        self.grad_power = torch.empty_like(power)
        self.stride = int(_native.call("wc_vjp_partial_stride", material_count))
        rows = int(_native.call("wc_vjp_partial_rows", inputs.tile_count))
        self.partials = torch.empty((rows, self.stride), dtype=torch.float32, device=device)
        self.grad_params = torch.empty(2 * material_count + 1, dtype=torch.float32, device=device)
        self.grad_tx = torch.empty((int(inputs.view.tx_count), 3), dtype=torch.float32, device=device)
        self.loss = torch.empty((), dtype=torch.float32, device=device)
        p = _native.ptr
        self.view = _native.WcAdjoint(
            grad_power_total=p(self.grad_power), grad_field_total=None, grad_component_power=None,
            grad_component_field=None, grad_diffraction_power=None, power_total=p(self.out.power_total),
            field_total=p(self.out.field_total), component_field=p(self.out.component_field),
            diffraction_vector=p(self.out.diffraction_vector), topology=p(self.out.topology),
            path_count=p(self.out.path_count), topology_capacity=self.out.topology_capacity,
            partials=p(self.partials),
        )
        self.graph: torch.cuda.CUDAGraph | None = None
        self.vertices = vertices
        if vertices:
            if dscene.edge_count and int(inputs.view.components) & 4:
                raise RuntimeError("vertex gradients of diffraction paths are not available; solve without "
                                   "diffraction to differentiate the mesh vertices")
            g, e = dscene.group_count, dscene.edge_count
            self.grad_plane = torch.zeros((g, 4), dtype=torch.float64, device=device)
            self.grad_edge = torch.zeros((e, 8), dtype=torch.float64, device=device)
            self.contrib = torch.empty(9 * g + 12 * e, dtype=torch.float32, device=device)
            self.grad_vertices = torch.empty((int(dscene.vertices.shape[0]), 3), dtype=torch.float32, device=device)
            self.csr = dscene.vertex_csr()

    def launch(self, reuse_visibility: bool = False) -> None:
        """One step. ``reuse_visibility``: the chain-validity and edge-visibility
        masks in the workspace are those of this solve's geometry (left by a
        step of the same transmitters and receivers at another carrier), so
        only the field passes run; the outputs are identical."""
        if reuse_visibility:
            engine.radiomap_pass(self.dscene, self.inputs, self.out, 2)
            if int(self.inputs.view.components) & 4:
                engine.radiomap_pass(self.dscene, self.inputs, self.out, 4)
        else:
            engine.radiomap_forward(self.dscene, self.inputs, topology_capacity=self.capacity, out=self.out)
        power = self.out.power_total
        torch.div(self.scale, torch.clamp_min(power, 1.0e-38), out=self.grad_power)
        self.adjoint()
        self.reduce()
        if self.vertices:
            self.geometry()

    def geometry(self) -> None:
        """Vertex gradient of the same adjoint (plane accumulators cleared by a
        fill kernel, then wc_radiomap_vjp_geometry and the per-vertex sums)."""
        d = self.dscene
        p = _native.ptr
        stream = _native.stream_handle(d.device)
        self.grad_plane.zero_()
        _native.call("wc_radiomap_vjp_geometry", d.ref(), ctypes.byref(self.inputs.view), ctypes.byref(self.view),
                     p(self.inputs.tiles), self.inputs.tile_count, p(self.grad_plane), p(self.grad_edge), stream)
        offset, slot = self.csr
        _native.call("wc_geometry_vertex_grad", p(d.vertices), p(d.faces), p(d.group_face), d.group_count,
                     p(self.grad_plane), p(d.edge_vertices), d.edge_count, p(self.grad_edge), p(offset), p(slot),
                     int(d.vertices.shape[0]), p(self.contrib), p(self.grad_vertices), stream)

    def adjoint(self, phase: int = 0) -> None:
        """The adjoint launches alone (per-tile partial gradients); phase 1 or 2
        runs only the line-of-sight/reflection/transmission or the
        diffraction launch."""
        inputs = self.inputs
        args = (self.dscene.ref(), ctypes.byref(inputs.view), ctypes.byref(self.view), _native.ptr(inputs.tiles),
                inputs.tile_count)
        stream = _native.stream_handle(self.dscene.device)
        if phase:
            _native.call("wc_radiomap_vjp_phase", *args, int(phase), stream)
        else:
            _native.call("wc_radiomap_vjp", *args, stream)

    def reduce(self) -> None:
        """Deterministic column sums of the per-tile partials."""
        inputs = self.inputs
        p = _native.ptr
        _native.call("wc_vjp_reduce", p(self.partials), inputs.tile_count, self.stride, p(inputs.tiles),
                     int(inputs.view.tx_count), p(self.grad_params), p(self.grad_tx),
                     _native.stream_handle(self.dscene.device))

    def capture(self) -> None:
        side = torch.cuda.Stream(self.dscene.device)
        side.wait_stream(torch.cuda.current_stream(self.dscene.device))
        with torch.cuda.stream(side):
            self.launch()
        torch.cuda.current_stream(self.dscene.device).wait_stream(side)
        torch.cuda.synchronize(self.dscene.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.launch()
        torch.cuda.synchronize(self.dscene.device)

    def step(self) -> None:
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()
