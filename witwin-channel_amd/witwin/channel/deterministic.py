"""Deterministic RF solver: ``Config``, ``PathTable``, ``Result``, ``solve``.

Drop-in surface of witwin/channel/deterministic.py:105-971. A solve compiles
the scene (cached per scene object), plans reflection chains per transmitter
and runs the fused gfx950 radiomap kernel, which discovers every path of every
(transmitter, cell) pair, evaluates its field and sums the cell coherently in
float64 registers — no per-path rows are materialized unless
``export_paths=True``. ``ad_mode="vjp"`` attaches the fixed-topology adjoint
and ``ad_mode="jvp"`` its forward-mode tangents (under
``torch.autograd.forward_ad``), both in
:mod:`witwin.channel.kernels.deterministic`.

Result layout follows the reference: ``path_gain``/``field`` are
``[T, cols, rows]`` for a single ``ReceiverGrid`` (the flat receiver index is
``row * cols + col``, reference propagation/geometry.py:72-89) and ``[T, R]``
for point receivers; ``component_power`` holds los/reflection/diffraction plus
transmission when requested; for a single grid the diffraction power is the
vector-coherent sum over edges (reference deterministic.py:870-882).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from witwin.channel import _native
from witwin.channel.components import (
    AD_MODES as _VALID_AD_MODES,
    DEFAULT_COMPONENTS as _DEFAULT_COMPONENTS,
    DEPTH_CAPPED_COMPONENTS,
    component_bits,
    component_max_depth,
    validate_max_depth,
    validated_components,
)
from witwin.channel.constants import PHASE_CONVENTION
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import CompiledScene, compile as compile_scene, device_scene

_VALID_SORT_KEYS = frozenset({"receiver_transmitter_depth_component"})
_VALID_MAX_PATHS_SCOPES = frozenset({"global", "per_pair"})
_BASE_COMPONENTS = ("los", "reflection", "diffraction")
_SLOTS = {"los": 0, "reflection": 1, "diffraction": 2, "transmission": 3}


@dataclass(frozen=True, slots=True)
class Config:
    max_depth: int = 1
    max_diffraction_order: int = 1
    components: frozenset[str] | set[str] | tuple[str, ...] | list[str] = _DEFAULT_COMPONENTS
    coherent: bool = True
    return_field: bool = True
    export_paths: bool = False
    max_paths: int | None = None
    max_paths_scope: str = "global"
    sort_key: str = "receiver_transmitter_depth_component"
    diagnostics: bool = False
    ad_mode: str = "none"
    coupled_paths: bool = False
    coupled_candidate_limit: int = 1_000_000
    scattering_samples_per_m2: float = 8.0
    scattering_max_paths_per_pair: int = 4096
    scattering_power_threshold: float = 0.0
    scattering_coherent: bool = False
    scattering_chain_max_depth: int = 0
    scattering_chain_samples_per_m2: float = 2.0
    scattering_chain_max_rows: int = 256
    isb_boundary_taper: bool = False
    isb_boundary_taper_width: float = 0.5

    def __post_init__(self) -> None:
        validate_max_depth(self.max_depth)
        if not (0.0 < self.isb_boundary_taper_width <= 4.0):
            raise ValueError("isb_boundary_taper_width must be in (0, 4]")
        if self.max_diffraction_order < 0:
            raise ValueError("max_diffraction_order must be 0 or 1")
        if self.max_diffraction_order > 1:
            raise RuntimeError("max_diffraction_order above 1 is not supported yet")
        components = validated_components(
            self.components, error_message="components must be a non-empty subset of {valid}")
        if self.scattering_coherent and "scattering" not in components:
            raise RuntimeError("scattering_coherent=True requires the 'scattering' component")
        if self.max_depth > 5 and components & DEPTH_CAPPED_COMPONENTS:
            raise RuntimeError("deterministic reflection/transmission currently support max_depth <= 5")
        if self.coupled_paths:
            if self.max_depth < 2:
                raise RuntimeError("coupled reflection-diffraction paths require max_depth >= 2")
            if not {"reflection", "diffraction"}.issubset(components):
                raise RuntimeError("coupled paths require both reflection and diffraction components")
        if self.coupled_candidate_limit <= 0:
            raise ValueError("coupled_candidate_limit must be positive")
        if self.max_paths is not None and self.max_paths <= 0:
            raise ValueError("max_paths must be positive when set")
        if self.max_paths_scope not in _VALID_MAX_PATHS_SCOPES:
            raise ValueError("max_paths_scope must be 'global' or 'per_pair'")
        if self.sort_key not in _VALID_SORT_KEYS:
            raise ValueError(f"sort_key must be one of {sorted(_VALID_SORT_KEYS)}")
        if self.ad_mode not in _VALID_AD_MODES:
            raise RuntimeError(f"deterministic ad_mode must be one of {sorted(_VALID_AD_MODES)}")
        object.__setattr__(self, "components", components)


@dataclass(frozen=True, slots=True)
class PathTable:
    valid: torch.Tensor
    tx_id: torch.Tensor
    rx_id: torch.Tensor
    depth: torch.Tensor
    component_id: torch.Tensor
    primitive_id: torch.Tensor
    edge_id: torch.Tensor
    path_length_m: torch.Tensor
    delay_s: torch.Tensor
    path_gain: torch.Tensor
    interaction_position: torch.Tensor
    interaction_normal: torch.Tensor
    material_id: torch.Tensor
    primitive_sequence: torch.Tensor
    material_sequence: torch.Tensor
    interaction_positions: torch.Tensor
    interaction_normals: torch.Tensor
    field_real: torch.Tensor
    field_imag: torch.Tensor
    coefficient: torch.Tensor
    field_xyz: torch.Tensor
    field_direction: torch.Tensor
    phase_rad: torch.Tensor
    interaction_count: torch.Tensor


@dataclass(frozen=True, slots=True)
class Result:
    path_gain: torch.Tensor
    field: torch.Tensor
    component_power: dict[str, torch.Tensor]
    component_fields: dict[str, torch.Tensor]
    paths: PathTable | None
    metadata: dict[str, Any]
    diagnostics: dict[str, Any] | None = None


def apply_layout(values: torch.Tensor, compiled: CompiledScene) -> torch.Tensor:
    ep = compiled.endpoints
    if ep.layout_kind == "grid":
        rows, cols = ep.grid_shape  # type: ignore[misc]
        return values.reshape(values.shape[0], rows, cols).transpose(1, 2).contiguous()
    return values.contiguous()


def phase_from_field(field: torch.Tensor) -> torch.Tensor:
    """(-arg F) mod 2 pi (reference deterministic.cu:390-406)."""
    phase = -torch.atan2(field.imag, field.real)
    phase = torch.remainder(phase, 2.0 * np.pi)
    return phase.to(torch.float32)


def build_path_table(rows: dict[str, Any], *, include_fields: bool = True) -> PathTable:
    path_field = rows["path_field"] if include_fields else torch.zeros_like(rows["path_field"])
    count = int(rows["tx_id"].shape[0])
    device = rows["tx_id"].device
    width = rows["primitive_sequence"].shape[1] if count else 1
    return PathTable(
        valid=torch.ones(count, dtype=torch.bool, device=device),
        tx_id=rows["tx_id"], rx_id=rows["rx_id"], depth=rows["depth"], component_id=rows["component_id"],
        primitive_id=rows["primitive_id"], edge_id=rows["edge_id"], path_length_m=rows["path_length_m"],
        delay_s=rows["delay_s"], path_gain=rows["path_gain"],
        interaction_position=rows["interaction_positions"][:, 0, :].contiguous() if count else
        torch.zeros((0, 3), device=device),
        interaction_normal=rows["interaction_normals"][:, 0, :].contiguous() if count else
        torch.zeros((0, 3), device=device),
        material_id=rows["material_id"], primitive_sequence=rows["primitive_sequence"].reshape(count, width),
        material_sequence=rows["material_sequence"].reshape(count, width),
        interaction_positions=rows["interaction_positions"].reshape(count, width, 3),
        interaction_normals=rows["interaction_normals"].reshape(count, width, 3),
        field_real=path_field.real.contiguous(), field_imag=path_field.imag.contiguous(),
        coefficient=rows["coefficient"], field_xyz=rows["field_xyz"], field_direction=rows["arrival_direction"],
        phase_rad=phase_from_field(path_field), interaction_count=rows["depth"],
    )


def accumulate_rows(rows: dict[str, Any], num_tx: int, num_rx: int, coherent: bool,
                    tx_power: torch.Tensor) -> tuple[torch.Tensor, ...]:
    """Per-cell sums of explicit rows (wc_accumulate_flat: slot map and
    coherent / incoherent finalize of reference deterministic.cu:829-970) plus
    the grid diffraction override's vector sum over the rows' field vectors
    (reference deterministic.py:870-882). Returns power_total [T*R],
    field_total [T*R] c64, component_power [4, T*R] (los, reflection,
    diffraction, transmission), component_field [4, T*R] c64 and the vector
    diffraction power [T*R]."""
    device = rows["tx_id"].device
    count = int(rows["tx_id"].shape[0])
    cells = num_tx * num_rx
    cp = torch.zeros((6, cells), dtype=torch.float32, device=device)
    cf = torch.zeros((6, cells), dtype=torch.complex64, device=device)
    pt = torch.zeros(cells, dtype=torch.float32, device=device)
    ft = torch.zeros(cells, dtype=torch.complex64, device=device)
    diff = torch.zeros(cells, dtype=torch.float32, device=device)
    if count:
        field = rows["path_field"].to(torch.complex64)
        args = (torch.ones(count, dtype=torch.uint8, device=device), rows["tx_id"].to(torch.int32).contiguous(),
                rows["rx_id"].to(torch.int32).contiguous(), rows["component_id"].to(torch.int32).contiguous(),
                rows["path_gain"].to(torch.float32).contiguous(), field.real.contiguous(), field.imag.contiguous())
        work = torch.empty(int(_native.call("wc_accumulate_workspace_bytes", count, num_tx, num_rx)),
                           dtype=torch.uint8, device=device)
        p = _native.ptr
        _native.call("wc_accumulate_flat", count, *[p(a) for a in args], num_tx, num_rx, 1 if coherent else 0,
                     p(cp), p(torch.view_as_real(cf)), p(pt), p(torch.view_as_real(ft)), p(work), work.numel(),
                     _native.stream_handle(device))
        is_diff = rows["component_id"] == 2
        if bool(is_diff.any()):
            cell = rows["tx_id"][is_diff].long() * num_rx + rows["rx_id"][is_diff].long()
            amp = tx_power.to(device)[rows["tx_id"][is_diff].long()].clamp_min(0.0).sqrt()
            vec = torch.zeros((cells, 3), dtype=torch.complex64, device=device)
            vec.index_add_(0, cell, rows["field_xyz"][is_diff] * amp[:, None])
            diff = vec.abs().square().sum(dim=-1)
    return pt, ft, cp[:4], cf[:4], diff


def _add_coupled(dscene, inputs, config: Config, power_total, field_total, component_power, component_fields,
                 num_tx: int, num_rx: int):
    """Coupled R->D / D->R / D->D paths into the dedicated coupled slot
    (reference deterministic.py:265-290, ADR-011 D2: component ids 3, 4 and 7
    accumulate coherently into one slot that joins the field sum). Coherent:
    field_total += coupled field, power_total = |field_total|^2; incoherent:
    power_total += coupled gains, field_total = sqrt(power_total)."""
    co = engine.coupled_fields(dscene, inputs, candidate_limit=config.coupled_candidate_limit)
    cf = co.field.reshape(num_tx, num_rx)
    if config.coherent:
        total = field_total.to(torch.complex128) + cf.to(torch.complex128)
        field_total = total.to(torch.complex64)
        power_total = (total.real.square() + total.imag.square()).to(torch.float32)
        component_power["coupled"] = (cf.real.double().square() + cf.imag.double().square()).float()
    else:
        power_total = (power_total.double() + co.power.reshape(num_tx, num_rx).double()).float()
        field_total = power_total.clamp_min(0.0).sqrt().to(torch.complex64)
        component_power["coupled"] = co.power.reshape(num_tx, num_rx)
    component_fields["coupled"] = cf
    total_n = int(co.count.sum().item())
    dd_n = int(co.count_dd.sum().item())
    return power_total, field_total, {"coupled": total_n, "coupled_double_diffraction": dd_n}


def _validate(config: Config) -> None:
    unsupported = set(config.components) - {"los", "reflection", "diffraction", "transmission"}
    if unsupported:
        raise NotImplementedError(f"components {sorted(unsupported)} are outside this build's hot path")
    if config.coupled_paths and config.ad_mode in ("vjp", "jvp"):
        raise NotImplementedError("coupled reflection-diffraction paths are forward only in this build "
                                  f"(ad_mode='{config.ad_mode}')")
    if config.coupled_paths and (config.export_paths or config.max_paths is not None):
        raise NotImplementedError("coupled reflection-diffraction paths enter the radiomap only; export_paths and "
                                  "max_paths do not carry coupled rows in this build")
    if config.isb_boundary_taper:
        raise NotImplementedError("isb_boundary_taper is outside this build's hot path")
    if config.ad_mode in ("vjp", "jvp") and config.max_paths is not None:
        raise NotImplementedError(f"ad_mode='{config.ad_mode}' differentiates every path of a cell; max_paths "
                                  "truncation is not supported with it")
    if "reflection" in config.components and config.max_depth < 1:
        raise RuntimeError("deterministic reflection requires max_depth >= 1")
    if "diffraction" in config.components:
        if config.max_depth < 1:
            raise RuntimeError("deterministic diffraction requires max_depth >= 1")
        if config.max_diffraction_order < 1:
            raise RuntimeError("deterministic diffraction requires max_diffraction_order >= 1")


def _metadata(config: Config, path_count: int, counts: dict[str, int], launches: int) -> dict[str, Any]:
    status = {name: ("enabled" if name in config.components else "not_requested")
              for name in ("los", "reflection", "diffraction", "transmission", "scattering")}
    if "transmission" in config.components and counts.get("transmission", 0) == 0:
        status["transmission"] = "enabled_no_paths"
    meta: dict[str, Any] = {
        "max_depth": config.max_depth,
        "max_diffraction_order": config.max_diffraction_order,
        "coherent": config.coherent,
        "return_field": config.return_field,
        "export_paths": config.export_paths,
        "max_paths": config.max_paths,
        "max_paths_scope": config.max_paths_scope,
        "sort_key": config.sort_key,
        "accumulation_strategy": "coherent" if config.coherent else "incoherent",
        "components": status,
        "counts": {"path_count": path_count, "valid_path_count": path_count, "components": counts},
        "capability": {"rayd_native": False, "path_native": True, "hip_native": True, "cuda_available": False,
                       "optix_available": False},
        "kernel": {"primitive": "deterministic_solver", "forward_launch_count": launches,
                   "accumulation_strategy": "per_cell_float64_registers", "scheduling_strategy": "fused_hip",
                   "ad_status": config.ad_mode},
        "field_abi": "complex3_v1",
        "phase_convention": dict(PHASE_CONVENTION),
        "coefficient_semantics": "unit_excitation_dimensionless_receiver_projection",
        "component_max_depth": component_max_depth(config.components, chain_depth=config.max_depth,
                                                   single_bounce_depth=1),
    }
    if "transmission" in config.components:
        meta["transmission"] = {"thin_sheet_straight_path_approximation": True, "group_delay": "geometric"}
    meta["coupled_paths"] = ({"requested": False, "geometry": "not_requested", "coefficient": "not_requested"}
                             if not config.coupled_paths else
                             {"requested": True, "geometry": "native_1r1d_reciprocal_plus_dd",
                              "coefficient": "unified_complex3_jones",
                              "double_diffraction": {"geometry": "native_dd_cascade", "component_id": 7,
                                                     "row_count": int(counts.get("coupled_double_diffraction", 0))}})
    return meta


def _solve_pipeline(compiled: CompiledScene, config: Config, *, tx_position: torch.Tensor | None = None,
                    material_params: Any = None) -> Result:
    _validate(config)
    device = _native.require_gpu()
    ep = compiled.endpoints
    if len(ep.rx_position) == 0 or len(ep.tx_position) == 0:
        raise RuntimeError("deterministic solve requires at least one transmitter and one receiver")
    dscene = device_scene(compiled, device)
    bits = component_bits(config.components)
    depth = config.max_depth if bits & (2 | 4 | 8) else 0
    if bits & 8:
        bits |= 16  # transmission penetration overflow fails loudly, as in the reference
    num_tx = int(ep.tx_position.shape[0])
    num_rx = int(ep.rx_position.shape[0])
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=config.coherent,
                                 frequency_hz=compiled.frequency_hz, tx_pos=tx_position)
    if config.ad_mode in ("vjp", "jvp"):
        from witwin.channel.kernels.deterministic import radiomap_with_vjp

        out, tensors = radiomap_with_vjp(compiled, dscene, inputs, material_params=material_params,
                                         tx_position=tx_position)
        power_total, field_total, comp_power, diff_power = tensors
    else:
        out = engine.radiomap_forward(dscene, inputs)
        power_total, field_total, comp_power = out.power_total, out.field_total, out.component_power
        diff_power = out.diffraction_vector.abs().square().sum(dim=-1)
    power_total = power_total.reshape(num_tx, num_rx)
    field_total = field_total.reshape(num_tx, num_rx)
    names = list(_BASE_COMPONENTS) + (["transmission"] if "transmission" in config.components else [])
    component_power = {name: comp_power[_SLOTS[name]].reshape(num_tx, num_rx) for name in names}
    component_fields = {name: out.component_field[_SLOTS[name]].reshape(num_tx, num_rx) for name in names}
    if ep.layout_kind == "grid" and "diffraction" in config.components:
        exact = diff_power.reshape(num_tx, num_rx)
        if not config.coherent:
            power_total = power_total - component_power["diffraction"] + exact
        component_power["diffraction"] = exact
    counts_t = out.path_count.reshape(num_tx, num_rx)
    path_count = int(counts_t.sum().item())
    coupled_counts = None
    if config.coupled_paths:
        power_total, field_total, coupled_counts = _add_coupled(dscene, inputs, config, power_total, field_total,
                                                                component_power, component_fields, num_tx, num_rx)
        path_count += coupled_counts["coupled"]
    paths = None
    counts = {"los": 0, "reflection": 0, "diffraction": 0}
    truncated = False
    if config.export_paths or config.max_paths is not None:
        per_pair = config.max_paths if config.max_paths_scope == "per_pair" else None
        rows = engine.export_rows(dscene, inputs, width=max(depth, 1), max_paths_per_pair=per_pair)
        truncated = rows["dropped"] > 0
        if config.max_paths is not None and config.max_paths_scope == "global":
            count = int(rows["tx_id"].shape[0])
            truncated = count > config.max_paths
            rows = {k: (v[:config.max_paths] if isinstance(v, torch.Tensor) and k not in ("pair_offset", "pair_cell")
                        else v) for k, v in rows.items()}
        if truncated:
            # the reference truncates before fields and accumulation
            # (propagation/enumerated.py:788-792): the maps sum the kept rows only
            power_total, field_total, comp_p4, comp_f4, diff_t = accumulate_rows(
                rows, num_tx, num_rx, config.coherent, inputs.tx_power)
            power_total = power_total.reshape(num_tx, num_rx)
            field_total = field_total.reshape(num_tx, num_rx)
            component_power = {name: comp_p4[_SLOTS[name]].reshape(num_tx, num_rx) for name in names}
            component_fields = {name: comp_f4[_SLOTS[name]].reshape(num_tx, num_rx) for name in names}
            if ep.layout_kind == "grid" and "diffraction" in config.components:
                exact = diff_t.reshape(num_tx, num_rx)
                if not config.coherent:
                    power_total = power_total - component_power["diffraction"] + exact
                component_power["diffraction"] = exact
        path_count = int(rows["tx_id"].shape[0])
        cid = rows["component_id"]
        for name, c in (("los", 0), ("reflection", 1), ("diffraction", 2), ("transmission", 5)):
            counts[name] = int((cid == c).sum().item())
        if config.export_paths:
            paths = build_path_table(rows, include_fields=config.return_field)
    if coupled_counts is not None:
        counts.update(coupled_counts)
    metadata = _metadata(config, path_count, counts, launches=2 if config.export_paths else 1)
    path_gain = apply_layout(power_total, compiled)
    field = apply_layout(field_total, compiled) if config.return_field else \
        torch.empty((0,), device=device, dtype=torch.complex64)
    comp_p = {k: apply_layout(v, compiled) for k, v in component_power.items()}
    comp_f = {k: (apply_layout(v, compiled) if config.return_field else
                  torch.empty((0,), device=device, dtype=torch.complex64)) for k, v in component_fields.items()}
    diagnostics = None
    if config.diagnostics:
        diagnostics = {
            "path_gain_shape": tuple(path_gain.shape), "field_shape": tuple(field.shape),
            "path_count": path_count, "component_counts": counts, "coherent": config.coherent,
            "accumulation_mode": "coherent" if config.coherent else "incoherent",
            "native_launch_count": metadata["kernel"]["forward_launch_count"],
            "diffraction_accumulation": ("exact_vector_coherent_paths" if ep.layout_kind == "grid"
                                         else "scalar_path_accumulation"),
            "path_planning": {"max_paths": config.max_paths, "max_paths_scope": config.max_paths_scope,
                              "candidate_count": inputs.chains.count, "truncated": truncated},
        }
    failure = int(out.failure.item())
    if failure & 4:
        # reference transmission.py: a segment crossing more walls than the
        # hit capacity (max_depth) fails loudly
        raise RuntimeError("deterministic transmission: a transmitter-receiver segment crosses more walls than "
                           f"max_depth={config.max_depth} (penetration overflow)")
    if failure:
        raise RuntimeError(f"deterministic solve exceeded a native capacity (failure bits {failure})")
    return Result(path_gain=path_gain, field=field, component_power=comp_p, component_fields=comp_f, paths=paths,
                  metadata=metadata, diagnostics=diagnostics)


_COMPILE_CACHE: dict[tuple[int, float], CompiledScene] = {}


def compile_cached(scene: Any, reference_frequency_hz: Any) -> CompiledScene:
    freq = float(reference_frequency_hz.detach()) if isinstance(reference_frequency_hz, torch.Tensor) \
        else float(reference_frequency_hz)
    key = (id(scene), freq)
    cached = _COMPILE_CACHE.get(key)
    if cached is None or cached.cache.get("scene_ref") is not scene:
        cached = compile_scene(scene, reference_frequency_hz=freq)
        cached.cache["scene_ref"] = scene
        _COMPILE_CACHE[key] = cached
    return cached


def solve(scene, config: Config, *, reference_frequency_hz) -> Result:  # type: ignore[no-untyped-def]
    """Run the deterministic propagation pipeline."""

    compiled = compile_scene(scene, reference_frequency_hz=reference_frequency_hz)
    return _solve_pipeline(compiled, config)
