"""Solver-neutral propagation consumer: explicit source and sink endpoint
batches against a compiled scene, compact path rows with exact pair
segmentation, and the scalar / complex3 / polarimetric (Jones) transports.

Mirrors the reference consumer contract (witwin/channel/propagation/
consumer.py): the vocabulary (:405-435), ``EndpointBatch`` (:515-588),
``PropagationRequest`` (:623-653), ``PropagationTopology`` /
``PropagationGeometry`` (:656-793), the transports (:796-916),
``PropagationPathBatch`` (:919-960), ``PropagationConvention`` (:963-1005),
``capabilities`` (:1181), ``evaluate`` (:3478-3520) and the fixed-topology
replay: ``prepare_fixed_topology`` (:1518), ``replicate_over_slots`` (:1573),
``FixedTopologyRequest`` (:1307), ``FixedTopologyEvaluation`` (:1421) and
``reevaluate`` (:3965-4110), with slot batching and wideband frequency
offsets.

The rows come from the same HIP engine as ``deterministic.solve`` with
``export_paths`` (count -> one count read -> emit -> canonical order), run on
the request's endpoints instead of the compiled scene's: rows are pair-major
with ``pair_index = sink_index * source_count + source_index``. The Jones
response is line-of-sight only, as in the reference; it excites the native
free-space row operator with the two source basis vectors and projects onto
the sink basis. With ``ad_mode`` ``"vjp"`` or ``"jvp"``, ``evaluate``
discovers natively and takes its line-of-sight and reflection rows' values
from the differentiable row operators (as ``reevaluate`` does). ``reevaluate`` replays
frozen line-of-sight and reflection rows at new endpoints in one native launch
per frequency column (``wc_fixed_rows_evaluate``); with ``ad_mode`` ``"vjp"``
or ``"jvp"`` its line-of-sight rows take their values from the
differentiable free-space row operator (reference consumer.py:3994-4017).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import torch

from witwin.channel import _native
from witwin.channel.components import component_bits
from witwin.channel.constants import NARROWBAND_FREQUENCY_OFFSET_LAW, PHASOR, TIME_DEPENDENCE
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import CompiledScene, device_scene

CONTRACT_VERSION = 6
COMPONENTS: frozenset[str] = frozenset({"los", "reflection", "transmission", "diffraction"})
RESPONSES: frozenset[str] = frozenset({"scalar_transport", "complex3_transport", "polarimetric_transport"})
TOPOLOGY_MODES: frozenset[str] = frozenset({"discover"})
AD_MODES: frozenset[str] = frozenset({"none", "jvp", "vjp"})
WORLD_MOTIONS: frozenset[str] = frozenset({"frozen_world", "fixed_winner_replay"})
# component ids a frozen topology may carry into a replay (reference consumer.py:211-214)
FIXED_TOPOLOGY_COMPONENT_IDS: tuple[tuple[str, int], ...] = (("los", 0), ("reflection", 1))
MAX_DEPTH = 5


def _require_tensor(name: str, value: Any, *, dtype: torch.dtype, shape: tuple[int, ...] | None = None,
                    ndim: int | None = None, device: torch.device | None = None) -> torch.Tensor:
    if not isinstance(value, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if value.dtype != dtype:
        raise TypeError(f"{name} must have dtype {dtype}")
    if ndim is not None and value.ndim != ndim:
        raise ValueError(f"{name} must have {ndim} dimensions")
    if shape is not None and tuple(value.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}")
    if device is not None and value.device != device:
        raise ValueError(f"{name} must share the endpoint device")
    return value


@dataclass(frozen=True, slots=True, eq=False)
class EndpointBatch:
    """Explicit point endpoints of one request (reference consumer.py:515-588).
    ``powers_w`` is required for sources and absent for sinks."""

    stable_ids: torch.Tensor
    positions_m: torch.Tensor
    polarizations: torch.Tensor
    polarization_basis: torch.Tensor | None = None
    powers_w: torch.Tensor | None = None

    def __post_init__(self) -> None:
        positions = _require_tensor("positions_m", self.positions_m, dtype=torch.float32, ndim=2)
        if tuple(positions.shape[1:]) != (3,):
            raise ValueError("positions_m must have shape (N, 3)")
        if positions.device.type != "cuda":
            raise ValueError("endpoint tensors must be CUDA tensors")
        rows, device = int(positions.shape[0]), positions.device
        _require_tensor("stable_ids", self.stable_ids, dtype=torch.int64, shape=(rows,), device=device)
        _require_tensor("polarizations", self.polarizations, dtype=torch.float32, shape=(rows, 3), device=device)
        if self.polarization_basis is not None:
            _require_tensor("polarization_basis", self.polarization_basis, dtype=torch.float32,
                            shape=(rows, 2, 3), device=device)
        if self.powers_w is not None:
            _require_tensor("powers_w", self.powers_w, dtype=torch.float32, shape=(rows,), device=device)
        for name in ("stable_ids", "positions_m", "polarizations", "polarization_basis", "powers_w"):
            value = getattr(self, name)
            if value is not None and not value.is_contiguous():
                raise ValueError(f"{name} must be contiguous")

    @property
    def count(self) -> int:
        return int(self.positions_m.shape[0])

    @property
    def device(self) -> torch.device:
        return self.positions_m.device


def _require_vocabulary(name: str, value: object, allowed: frozenset[str]) -> str:
    if not isinstance(value, str):
        raise TypeError(f"{name} must be a string")
    if value not in allowed:
        raise NotImplementedError(f"unsupported {name} {value!r}; supported values are {sorted(allowed)}")
    return value


@dataclass(frozen=True, slots=True)
class PropagationRequest:
    """A discovery request against a compiled scene (reference consumer.py:623-653)."""

    sources: EndpointBatch
    sinks: EndpointBatch
    reference_frequency_hz: float | torch.Tensor
    components: frozenset[str]
    max_depth: int
    response: str
    topology_mode: str
    ad_mode: str
    max_paths: int | None = None

    def __post_init__(self) -> None:
        if not isinstance(self.sources, EndpointBatch) or not isinstance(self.sinks, EndpointBatch):
            raise TypeError("sources and sinks must be EndpointBatch instances")
        if self.sources.powers_w is None:
            raise ValueError("sources.powers_w is required")
        if self.sinks.powers_w is not None:
            raise ValueError("sinks.powers_w must be absent")
        if self.sources.device != self.sinks.device:
            raise ValueError("source and sink batches must share one CUDA device")
        if type(self.components) is not frozenset or not self.components:
            raise TypeError("components must be a non-empty frozenset")
        unsupported = self.components - COMPONENTS
        if unsupported:
            raise NotImplementedError(f"unsupported propagation components: {sorted(unsupported)}")
        _require_vocabulary("response", self.response, RESPONSES)
        _require_vocabulary("topology_mode", self.topology_mode, TOPOLOGY_MODES)
        _require_vocabulary("ad_mode", self.ad_mode, AD_MODES)
        if type(self.max_depth) is not int or not 0 <= self.max_depth <= MAX_DEPTH:
            raise ValueError(f"max_depth must be an int in [0, {MAX_DEPTH}]")
        if self.max_paths is not None and (type(self.max_paths) is not int or self.max_paths <= 0):
            raise ValueError("max_paths must be a positive int when set")


@dataclass(frozen=True, slots=True, eq=False)
class PropagationTopology:
    """Compact discrete rows in stable pair-major order."""

    source_index: torch.Tensor
    sink_index: torch.Tensor
    source_id: torch.Tensor
    sink_id: torch.Tensor
    depth: torch.Tensor
    component_id: torch.Tensor
    primitive_id: torch.Tensor
    edge_id: torch.Tensor
    material_id: torch.Tensor
    primitive_sequence: torch.Tensor
    material_sequence: torch.Tensor
    interaction_type: torch.Tensor
    provenance: Any = None

    @property
    def row_count(self) -> int:
        return int(self.source_index.shape[0])

    @property
    def device(self) -> torch.device:
        return self.source_index.device


@dataclass(frozen=True, slots=True, eq=False)
class PropagationGeometry:
    """Continuous SI geometry aligned with one compact topology."""

    path_length_m: torch.Tensor
    delay_s: torch.Tensor
    field_direction: torch.Tensor
    interaction_positions_m: torch.Tensor
    interaction_normals: torch.Tensor

    @property
    def row_count(self) -> int:
        return int(self.path_length_m.shape[0])

    @property
    def device(self) -> torch.device:
        return self.path_length_m.device


@dataclass(frozen=True, slots=True, eq=False)
class ScalarTransport:
    """Endpoint-projected complex coefficient including sqrt(source power)."""

    coefficient: torch.Tensor
    coefficient_offsets: torch.Tensor | None = None
    frequency_offsets_hz: tuple[float, ...] | None = None

    @property
    def row_count(self) -> int:
        return int(self.coefficient.shape[0])

    @property
    def device(self) -> torch.device:
        return self.coefficient.device


@dataclass(frozen=True, slots=True, eq=False)
class Complex3Transport:
    """World-Cartesian field including sqrt(source power), and its propagation
    direction at the sink."""

    field: torch.Tensor
    direction: torch.Tensor
    field_offsets: torch.Tensor | None = None
    frequency_offsets_hz: tuple[float, ...] | None = None

    @property
    def row_count(self) -> int:
        return int(self.field.shape[0])

    @property
    def device(self) -> torch.device:
        return self.field.device


@dataclass(frozen=True, slots=True, eq=False)
class JonesTransport:
    """Source-basis to sink-basis 2 x 2 operator (power excluded):
    ``matrix[k, i, j]`` is the response of sink basis vector i to source basis
    vector j."""

    matrix: torch.Tensor
    source_basis: torch.Tensor
    sink_basis: torch.Tensor

    @property
    def row_count(self) -> int:
        return int(self.matrix.shape[0])

    @property
    def device(self) -> torch.device:
        return self.matrix.device


@dataclass(frozen=True, slots=True, eq=False)
class PropagationPathBatch:
    """Compact ``K`` rows with their pair segmentation."""

    pair_count: int
    path_count: int
    pair_index: torch.Tensor
    pair_offsets: torch.Tensor
    topology: PropagationTopology
    geometry: PropagationGeometry
    transport: ScalarTransport | Complex3Transport | JonesTransport

    def __post_init__(self) -> None:
        if type(self.pair_count) is not int or self.pair_count < 0:
            raise ValueError("pair_count must be a non-negative int")
        if type(self.path_count) is not int or self.path_count < 0:
            raise ValueError("path_count must be a non-negative int")
        _require_tensor("pair_index", self.pair_index, dtype=torch.int64, shape=(self.path_count,))
        _require_tensor("pair_offsets", self.pair_offsets, dtype=torch.int64, shape=(self.pair_count + 1,))
        for part in (self.topology, self.geometry, self.transport):
            if part.row_count != self.path_count:
                raise ValueError("topology, geometry and transport must have exactly path_count rows")


@dataclass(frozen=True, slots=True)
class PropagationConvention:
    contract_version: int = CONTRACT_VERSION
    pair_layout: str = "sink_major_source_minor:pair_index=sink_index*source_count+source_index"
    # block-diagonal law of a slot-batched replay (reference consumer.py:975-983)
    slot_pair_layout: str = ("block_diagonal_slots:row=slot*frozen_row_count+frozen_row;"
                             "source_index=slot*slot_source_count+slot_source_index;"
                             "sink_index=slot*slot_sink_count+slot_sink_index;"
                             "pair_index=slot*slot_source_count*slot_sink_count"
                             "+slot_sink_index*slot_source_count+slot_source_index;"
                             "pair_count=slot_count*slot_source_count*slot_sink_count")
    distance_unit: str = "m"
    delay_unit: str = "s"
    phasor: str = PHASOR
    time_dependence: str = TIME_DEPENDENCE
    coefficient_reference: str = "includes_reference_frequency_phase"
    narrowband_frequency_offset_law: str = NARROWBAND_FREQUENCY_OFFSET_LAW
    complex3_basis: str = "world_cartesian"
    jones_mapping: str = "source_transverse_basis_to_sink_transverse_basis"


@dataclass(frozen=True, slots=True)
class PropagationCapabilities:
    contract_version: int
    components: frozenset[str]
    responses: frozenset[str]
    topology_modes: frozenset[str]
    ad_modes: frozenset[str]
    response_components: tuple[tuple[str, frozenset[str]], ...]
    response_ad_modes: tuple[tuple[str, frozenset[str]], ...]
    supports_los_jones: bool = True
    supports_fixed_topology: bool = True
    fixed_topology_components: frozenset[str] = frozenset({"los", "reflection"})
    fixed_topology_responses: frozenset[str] = frozenset(
        {"scalar_transport", "complex3_transport", "polarimetric_transport"})
    fixed_topology_row_validity_components: frozenset[str] = frozenset({"los", "reflection"})
    wideband_responses: frozenset[str] = frozenset({"scalar_transport", "complex3_transport"})
    world_motions: frozenset[str] = WORLD_MOTIONS
    # reevaluate differentiates line-of-sight and reflection rows (scalar and
    # complex3 transports) through the free-space and reflection row operators
    fixed_topology_ad_modes: frozenset[str] = frozenset({"none", "jvp", "vjp"})
    fixed_topology_ad_components: frozenset[str] = frozenset({"los", "reflection"})

    def components_for(self, response: str) -> frozenset[str]:
        return dict(self.response_components)[response]

    def ad_modes_for(self, response: str) -> frozenset[str]:
        return dict(self.response_ad_modes)[response]


_CAPABILITIES = PropagationCapabilities(
    contract_version=CONTRACT_VERSION,
    components=COMPONENTS,
    responses=RESPONSES,
    topology_modes=TOPOLOGY_MODES,
    ad_modes=frozenset({"none", "jvp", "vjp"}),
    response_components=(("scalar_transport", COMPONENTS), ("complex3_transport", COMPONENTS),
                         ("polarimetric_transport", frozenset({"los"}))),
    response_ad_modes=(("scalar_transport", AD_MODES), ("complex3_transport", AD_MODES),
                       ("polarimetric_transport", frozenset({"none"}))),
)


def capabilities() -> PropagationCapabilities:
    """What this consumer contract version supports (frozen record)."""
    return _CAPABILITIES


@dataclass(frozen=True, slots=True)
class PropagationDiagnostics:
    discovery_launch_count: int
    candidate_count: int
    visibility_rejection_count: int
    compact_count_d2h_copies: int
    compact_count_d2h_bytes: int
    compact_sync_count: int
    validation_d2h_copies: int = 0
    validation_d2h_bytes: int = 0
    validation_sync_count: int = 0
    frequency_column_count: int = 1
    ad_companion_launches: int = 0
    ad_tape_bytes: int = 0


@dataclass(frozen=True, slots=True, eq=False)
class PropagationEvaluation:
    paths: PropagationPathBatch
    convention: PropagationConvention = field(default_factory=PropagationConvention)
    capabilities: PropagationCapabilities = _CAPABILITIES
    diagnostics: PropagationDiagnostics | None = None


def _preflight(compiled: object, request: object) -> tuple[CompiledScene, PropagationRequest]:
    if not isinstance(compiled, CompiledScene):
        raise TypeError("evaluate requires a CompiledScene")
    if not isinstance(request, PropagationRequest):
        raise TypeError("request must be a PropagationRequest")
    allowed = _CAPABILITIES.components_for(request.response)
    if not request.components.issubset(allowed):
        raise NotImplementedError(f"{request.response} does not support components "
                                  f"{sorted(request.components - allowed)}")
    if request.ad_mode not in _CAPABILITIES.ad_modes_for(request.response):
        raise NotImplementedError(f"{request.response} does not support AD mode {request.ad_mode!r} in this "
                                  "build (differentiate through deterministic.solve(ad_mode='vjp'))")
    if request.ad_mode != "none" and not request.components <= _CAPABILITIES.fixed_topology_ad_components:
        raise NotImplementedError(f"ad_mode {request.ad_mode!r} differentiates line-of-sight and reflection rows; "
                                  f"the request asks for "
                                  f"{sorted(request.components - _CAPABILITIES.fixed_topology_ad_components)}")
    if request.response == "polarimetric_transport" and (request.sources.polarization_basis is None
                                                          or request.sinks.polarization_basis is None):
        raise ValueError("polarimetric_transport requires source and sink polarization_basis tensors")
    frequency = float(request.reference_frequency_hz.detach()) if isinstance(request.reference_frequency_hz,
                                                                             torch.Tensor) \
        else float(request.reference_frequency_hz)
    if frequency != compiled.frequency_hz:
        raise ValueError(f"reference_frequency_hz {frequency} does not match the compiled scene's "
                         f"{compiled.frequency_hz}")
    return compiled, request


def transverse_basis(reference: torch.Tensor, direction: torch.Tensor) -> torch.Tensor:
    """Row-aligned orthonormal basis (u, v) transverse to ``direction`` [K,3]
    from a reference pair [K,2,3]: u is the first reference vector with its
    component along the direction removed (Gram-Schmidt), v = d x u, flipped
    to point along the second reference vector. A first vector parallel to
    the direction falls back to the second one."""
    d = direction / torch.clamp(torch.linalg.vector_norm(direction, dim=-1, keepdim=True), min=1e-30)
    first, second = reference[:, 0], reference[:, 1]

    def perp(v: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        w = v - (v * d).sum(-1, keepdim=True) * d
        n = torch.linalg.vector_norm(w, dim=-1, keepdim=True)
        return w / torch.clamp(n, min=1e-30), n

    u1, n1 = perp(first)
    u2, _ = perp(second)
    u = torch.where(n1 > 1e-6, u1, u2)
    v = torch.cross(d, u, dim=-1)
    v = torch.where((v * second).sum(-1, keepdim=True) < 0, -v, v)
    return torch.stack((u, v), dim=1).contiguous()


def _free_space_field(source: torch.Tensor, target: torch.Tensor, polarization: torch.Tensor,
                      frequency_hz: float) -> tuple[torch.Tensor, torch.Tensor]:
    """Unit-power free-space complex3 field of each row (native row operator)."""
    n = int(source.shape[0])
    device = source.device
    out = {name: torch.empty(shape, device=device, dtype=dt) for name, shape, dt in (
        ("field_vector", (n, 3), torch.complex64), ("coefficient", (n,), torch.complex64),
        ("path_field", (n,), torch.complex64), ("path_gain", (n,), torch.float32),
        ("path_length", (n,), torch.float32), ("delay", (n,), torch.float32), ("direction", (n, 3), torch.float32))}
    power = torch.ones(n, device=device, dtype=torch.float32)
    p = _native.ptr
    if n:
        _native.call("wc_field_free_space", n, p(source), p(target), p(power), p(polarization), p(polarization),
                     float(frequency_hz), p(torch.view_as_real(out["field_vector"])),
                     p(torch.view_as_real(out["coefficient"])), p(torch.view_as_real(out["path_field"])),
                     p(out["path_gain"]), p(out["path_length"]), p(out["delay"]), p(out["direction"]),
                     _native.stream_handle(device))
    return out["field_vector"], out["direction"]


def _project(field_vector: torch.Tensor, direction: torch.Tensor, sink_vector: torch.Tensor) -> torch.Tensor:
    n = int(field_vector.shape[0])
    coefficient = torch.empty(n, device=field_vector.device, dtype=torch.complex64)
    gain = torch.empty(n, device=field_vector.device, dtype=torch.float32)
    p = _native.ptr
    if n:
        _native.call("wc_field_project_complex3", n, p(torch.view_as_real(field_vector.contiguous())),
                     p(direction.contiguous()), p(sink_vector.contiguous()), p(torch.view_as_real(coefficient)),
                     p(gain), _native.stream_handle(field_vector.device))
    return coefficient


def _los_jones(source: torch.Tensor, target: torch.Tensor, source_reference: torch.Tensor,
               sink_reference: torch.Tensor, frequency_hz: float) -> JonesTransport:
    direction = target - source
    source_basis = transverse_basis(source_reference, direction)
    sink_basis = transverse_basis(sink_reference, direction)
    columns = [_free_space_field(source, target, source_basis[:, j].contiguous(), frequency_hz) for j in (0, 1)]
    arrival = columns[0][1]
    matrix = torch.stack([torch.stack([_project(columns[j][0], arrival, sink_basis[:, i].contiguous())
                                       for j in (0, 1)], dim=-1) for i in (0, 1)], dim=-2)
    return JonesTransport(matrix=matrix.contiguous(), source_basis=source_basis, sink_basis=sink_basis)


def evaluate(compiled_scene: CompiledScene, request: PropagationRequest) -> PropagationEvaluation:
    """Discover and evaluate one compact propagation batch (reference
    consumer.py:3478-3520): every source x sink pair, all requested
    components, rows in canonical pair-major order."""

    compiled, request = _preflight(compiled_scene, request)
    device = _native.require_gpu()
    sources, sinks = request.sources, request.sinks
    dscene = device_scene(compiled, device)
    bits = component_bits(request.components)
    depth = request.max_depth if bits & (2 | 4 | 8) else 0
    if depth == 0:
        bits &= ~(2 | 4)      # reflection and diffraction need at least one interaction
    num_tx, num_rx = sources.count, sinks.count
    pair_count = num_tx * num_rx
    width = max(depth, 1)
    candidates = 0
    if pair_count:
        inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                     frequency_hz=compiled.frequency_hz, tx_pos=sources.positions_m,
                                     tx_pol=sources.polarizations.detach().cpu().numpy(),
                                     tx_power=sources.powers_w.detach().cpu().numpy(),
                                     rx_pos=sinks.positions_m.detach().cpu().numpy(),
                                     rx_pol=sinks.polarizations.detach().cpu().numpy())
        rows = engine.export_rows(dscene, inputs, width=width)
        for extra in ("pair_offset", "pair_cell", "dropped"):
            rows.pop(extra)
        # (source, sink) candidates: the straight segment, every reflection
        # chain of the source and every edge
        candidates = pair_count * (1 + (int(dscene.view.edge_count) if bits & 4 else 0))
        candidates += num_rx * int(inputs.chains.count)
    else:
        empty_i = torch.zeros(0, dtype=torch.int32, device=device)
        rows = {name: empty_i for name in ("tx_id", "rx_id", "depth", "component_id", "primitive_id", "edge_id",
                                           "material_id")}
        rows.update({"primitive_sequence": torch.zeros((0, width), dtype=torch.int32, device=device),
                     "material_sequence": torch.zeros((0, width), dtype=torch.int32, device=device),
                     "interaction_positions": torch.zeros((0, width, 3), device=device),
                     "interaction_normals": torch.zeros((0, width, 3), device=device),
                     "path_length_m": torch.zeros(0, device=device), "delay_s": torch.zeros(0, device=device),
                     "arrival_direction": torch.zeros((0, 3), device=device),
                     "path_field": torch.zeros(0, dtype=torch.complex64, device=device),
                     "field_xyz": torch.zeros((0, 3), dtype=torch.complex64, device=device)})
    from witwin.channel.path import interaction_types

    k = int(rows["tx_id"].shape[0])
    if request.max_paths is not None and k > request.max_paths:
        rows = {name: (v[:request.max_paths] if isinstance(v, torch.Tensor) else v) for name, v in rows.items()}
        k = request.max_paths
    source_index = rows["tx_id"].to(torch.int32)
    sink_index = rows["rx_id"].to(torch.int32)
    pair_index = sink_index.long() * num_tx + source_index.long()
    counts = torch.bincount(pair_index, minlength=pair_count) if pair_count else torch.zeros(0, dtype=torch.int64,
                                                                                             device=device)
    pair_offsets = torch.zeros(pair_count + 1, dtype=torch.int64, device=device)
    if pair_count:
        pair_offsets[1:] = torch.cumsum(counts, 0)
    comp = rows["component_id"].to(torch.int32)
    topology = PropagationTopology(
        source_index=source_index, sink_index=sink_index,
        source_id=sources.stable_ids[source_index.long()], sink_id=sinks.stable_ids[sink_index.long()],
        depth=rows["depth"].to(torch.int32), component_id=comp, primitive_id=rows["primitive_id"].to(torch.int32),
        edge_id=rows["edge_id"].to(torch.int32), material_id=rows["material_id"].to(torch.int32),
        primitive_sequence=rows["primitive_sequence"].reshape(k, width).to(torch.int32),
        material_sequence=rows["material_sequence"].reshape(k, width).to(torch.int32),
        interaction_type=interaction_types(comp, rows["depth"], width))
    if request.ad_mode != "none" and k:
        # discovery fixed the rows; their values come from the differentiable
        # row operators at the request's (possibly tangent-carrying) endpoints
        src_rows, snk_rows = source_index.long(), sink_index.long()
        values = _rows_ad({"path_field": rows["path_field"].to(torch.complex64),
                           "field_xyz": rows["field_xyz"].to(torch.complex64),
                           "path_length": rows["path_length_m"].to(torch.float32),
                           "delay": rows["delay_s"].to(torch.float32),
                           "direction": rows["arrival_direction"].to(torch.float32)},
                          torch.ones(k, dtype=torch.bool, device=device), topology, compiled,
                          sources.positions_m[src_rows].contiguous(), sinks.positions_m[snk_rows].contiguous(),
                          sources.powers_w[src_rows].contiguous(), sources.polarizations[src_rows].contiguous(),
                          sinks.polarizations[snk_rows].contiguous(), request.reference_frequency_hz,
                          compiled.frequency_hz)
        rows = {**rows, "path_field": values["path_field"], "field_xyz": values["field_xyz"],
                "path_length_m": values["path_length"], "delay_s": values["delay"],
                "arrival_direction": values["direction"]}
    geometry = PropagationGeometry(
        path_length_m=rows["path_length_m"].to(torch.float32), delay_s=rows["delay_s"].to(torch.float32),
        field_direction=rows["arrival_direction"].to(torch.float32),
        interaction_positions_m=rows["interaction_positions"].reshape(k, width, 3).to(torch.float32),
        interaction_normals=rows["interaction_normals"].reshape(k, width, 3).to(torch.float32))
    if request.response == "scalar_transport":
        transport: Any = ScalarTransport(coefficient=rows["path_field"].to(torch.complex64))
    elif request.response == "complex3_transport":
        amplitude = torch.sqrt(torch.clamp(sources.powers_w[source_index.long()], min=0.0))
        transport = Complex3Transport(field=(rows["field_xyz"] * amplitude.unsqueeze(-1)).to(torch.complex64),
                                      direction=geometry.field_direction)
    else:
        transport = _los_jones(sources.positions_m[source_index.long()], sinks.positions_m[sink_index.long()],
                               sources.polarization_basis[source_index.long()],
                               sinks.polarization_basis[sink_index.long()], compiled.frequency_hz)
    diagnostics = PropagationDiagnostics(
        discovery_launch_count=3 if pair_count else 0, candidate_count=candidates,
        visibility_rejection_count=max(candidates - k, 0), compact_count_d2h_copies=1 if pair_count else 0,
        compact_count_d2h_bytes=4 * pair_count, compact_sync_count=1 if pair_count else 0)
    paths = PropagationPathBatch(pair_count=pair_count, path_count=k, pair_index=pair_index.contiguous(),
                                 pair_offsets=pair_offsets, topology=topology, geometry=geometry,
                                 transport=transport)
    return PropagationEvaluation(paths=paths, diagnostics=diagnostics)


# ---------------------------------------------------------------- fixed-topology replay

@dataclass(frozen=True, slots=True, eq=False)
class FixedTopologyBucket:
    """One ``(component, depth)`` partition of a frozen topology: ascending
    row indices, so a bucket keeps frozen row order (reference
    consumer.py:1228-1249)."""

    component: str
    depth: int
    rows: torch.Tensor

    def __post_init__(self) -> None:
        _require_vocabulary("component", self.component, COMPONENTS)
        if type(self.depth) is not int or not 0 <= self.depth <= MAX_DEPTH:
            raise ValueError(f"depth must be an int in [0, {MAX_DEPTH}]")
        _require_tensor("rows", self.rows, dtype=torch.int64)

    @property
    def row_count(self) -> int:
        return int(self.rows.shape[0])


@dataclass(frozen=True, slots=True, eq=False)
class PreparedFixedTopology:
    """A frozen topology partitioned once for repeated replay (reference
    consumer.py:1252-1287). The counters describe the preparation itself."""

    topology: PropagationTopology
    buckets: tuple[FixedTopologyBucket, ...]
    prepare_d2h_copies: int
    prepare_d2h_bytes: int
    prepare_synchronizations: int
    slot_count: int = 1
    slot_source_count: int = 0
    slot_sink_count: int = 0

    @property
    def row_count(self) -> int:
        return self.topology.row_count

    @property
    def device(self) -> torch.device:
        return self.topology.primitive_sequence.device

    @property
    def provenance(self) -> Any:
        return self.topology.provenance


def _component_name(component_id: int) -> str:
    for name, cid in FIXED_TOPOLOGY_COMPONENT_IDS:
        if cid == component_id:
            return name
    raise NotImplementedError(
        f"fixed-topology reevaluation supports components "
        f"{sorted(_CAPABILITIES.fixed_topology_components)}; the frozen topology carries component id "
        f"{component_id}")


def _malformed_rows(topology: PropagationTopology, width: int) -> torch.Tensor:
    depth = topology.depth.to(torch.int64)
    bad = (depth < 0) | (depth > width)
    if width:
        slots = torch.arange(width, device=depth.device).unsqueeze(0)
        used = slots < depth.unsqueeze(1)
        seq = topology.primitive_sequence
        bad = bad | ((seq >= 0) != used).any(dim=1)
    return bad


def prepare_fixed_topology(topology: PropagationTopology) -> PreparedFixedTopology:
    """Partition a frozen topology by (component, depth) once (reference
    consumer.py:1518-1570). Synchronizes; call it once per frozen topology."""
    if not isinstance(topology, PropagationTopology):
        raise TypeError("topology must be a PropagationTopology")
    if topology.row_count == 0:
        return PreparedFixedTopology(topology=topology, buckets=(), prepare_d2h_copies=0, prepare_d2h_bytes=0,
                                     prepare_synchronizations=0)
    width = int(topology.primitive_sequence.shape[1])
    if bool(_malformed_rows(topology, width).any().item()):
        raise ValueError("frozen topology rows disagree with their interaction sequence padding; depth must be "
                         "in [0, sequence width] and unused slots must hold -1")
    key = topology.component_id.to(torch.int64) * (width + 1) + topology.depth.to(torch.int64)
    distinct = torch.unique(key).tolist()
    buckets = []
    for value in distinct:
        component_id, depth = divmod(int(value), width + 1)
        component = _component_name(component_id)
        if (component == "los") != (depth == 0):
            raise ValueError(f"a {component} row cannot have depth {depth}")
        buckets.append(FixedTopologyBucket(component=component, depth=depth,
                                           rows=torch.nonzero(key == value, as_tuple=False).reshape(-1)))
    return PreparedFixedTopology(topology=topology, buckets=tuple(buckets), prepare_d2h_copies=2 + len(buckets),
                                 prepare_d2h_bytes=1 + 8 * (len(distinct) + len(buckets)),
                                 prepare_synchronizations=2 + len(buckets))


def _tile(t: torch.Tensor, slot_count: int) -> torch.Tensor:
    return t.repeat((slot_count,) + (1,) * (t.ndim - 1))


def replicate_over_slots(prepared: PreparedFixedTopology, slot_count: int, *, source_count: int,
                         sink_count: int) -> PreparedFixedTopology:
    """Tile a prepared topology over ``slot_count`` block-diagonal slots
    (reference consumer.py:1573-1660): row ``t*K + r`` is frozen row ``r`` in
    slot ``t``, its source and sink indices shifted by ``t`` times the
    per-slot counts. Index arithmetic only; ``slot_count == 1`` returns the
    handle unchanged."""
    if not isinstance(prepared, PreparedFixedTopology):
        raise TypeError("prepared must be a PreparedFixedTopology")
    if type(slot_count) is not int or slot_count <= 0:
        raise ValueError("slot_count must be a positive int")
    for name, value in (("source_count", source_count), ("sink_count", sink_count)):
        if type(value) is not int or value <= 0:
            raise ValueError(f"{name} must be a positive int")
    if slot_count == 1:
        return prepared
    if prepared.slot_count != 1:
        raise ValueError("the topology is already replicated over slots")
    topo = prepared.topology
    k = topo.row_count
    device = topo.primitive_sequence.device
    slot = torch.arange(slot_count, device=device, dtype=torch.int32).repeat_interleave(k)
    source_index = _tile(topo.source_index, slot_count) + slot * source_count
    sink_index = _tile(topo.sink_index, slot_count) + slot * sink_count
    replicated = PropagationTopology(
        source_index=source_index.to(torch.int32), sink_index=sink_index.to(torch.int32),
        source_id=_tile(topo.source_id, slot_count), sink_id=_tile(topo.sink_id, slot_count),
        depth=_tile(topo.depth, slot_count), component_id=_tile(topo.component_id, slot_count),
        primitive_id=_tile(topo.primitive_id, slot_count), edge_id=_tile(topo.edge_id, slot_count),
        material_id=_tile(topo.material_id, slot_count),
        primitive_sequence=_tile(topo.primitive_sequence, slot_count),
        material_sequence=_tile(topo.material_sequence, slot_count),
        interaction_type=_tile(topo.interaction_type, slot_count), provenance=topo.provenance)
    offsets = torch.arange(slot_count, device=device, dtype=torch.int64) * k
    buckets = tuple(FixedTopologyBucket(component=b.component, depth=b.depth,
                                        rows=(b.rows.unsqueeze(0) + offsets.unsqueeze(1)).reshape(-1))
                    for b in prepared.buckets)
    return PreparedFixedTopology(topology=replicated, buckets=buckets, prepare_d2h_copies=0, prepare_d2h_bytes=0,
                                 prepare_synchronizations=0, slot_count=slot_count, slot_source_count=source_count,
                                 slot_sink_count=sink_count)


def _require_frequency_offsets(offsets: Any) -> tuple[float, ...] | None:
    if offsets is None:
        return None
    if not isinstance(offsets, tuple) or not offsets:
        raise ValueError("frequency_offsets_hz must be a non-empty tuple of floats or None")
    values = []
    for o in offsets:
        if isinstance(o, bool) or not isinstance(o, (int, float)) or not (abs(float(o)) < float("inf")):
            raise ValueError("frequency_offsets_hz entries must be finite floats")
        values.append(float(o))
    return tuple(values)


@dataclass(frozen=True, slots=True, eq=False)
class FixedTopologyRequest:
    """A replay request against an already-discovered topology (reference
    consumer.py:1307-1418). ``topology`` is a raw :class:`PropagationTopology`
    for the zero-interaction line-of-sight route, or a
    :class:`PreparedFixedTopology` for any route with interactions."""

    sources: EndpointBatch
    sinks: EndpointBatch
    reference_frequency_hz: float | torch.Tensor
    topology: PropagationTopology | PreparedFixedTopology
    response: str
    ad_mode: str
    world_motion: str = "frozen_world"
    slot_count: int = 1
    frequency_offsets_hz: tuple[float, ...] | None = None

    def __post_init__(self) -> None:
        if not isinstance(self.sources, EndpointBatch) or not isinstance(self.sinks, EndpointBatch):
            raise TypeError("sources and sinks must be EndpointBatch instances")
        if type(self.slot_count) is not int or self.slot_count <= 0:
            raise ValueError("slot_count must be a positive int")
        if self.slot_count > 1:
            if isinstance(self.topology, PropagationTopology):
                raise NotImplementedError("slot_count > 1 requires a PreparedFixedTopology; call "
                                          "prepare_fixed_topology first")
            for name, count in (("sources", self.sources.count), ("sinks", self.sinks.count),
                                ("topology rows", self.frozen_topology.row_count)):
                if count % self.slot_count:
                    raise ValueError(f"{name} carries {count} entries, which is not a multiple of "
                                     f"slot_count={self.slot_count}; every slot must hold the same number")
        object.__setattr__(self, "frequency_offsets_hz", _require_frequency_offsets(self.frequency_offsets_hz))
        if not isinstance(self.topology, (PropagationTopology, PreparedFixedTopology)):
            raise TypeError("topology must be a PropagationTopology or a PreparedFixedTopology")
        response = _require_vocabulary("response", self.response, RESPONSES)
        _require_vocabulary("ad_mode", self.ad_mode, AD_MODES)
        _require_vocabulary("world_motion", self.world_motion, WORLD_MOTIONS)
        if response not in _CAPABILITIES.fixed_topology_responses:
            raise NotImplementedError(f"fixed-topology reevaluation does not support {response!r}")
        if self.frequency_offsets_hz is not None and response not in _CAPABILITIES.wideband_responses:
            raise NotImplementedError(f"frequency_offsets_hz is not supported for {response!r}; "
                                      f"capabilities().wideband_responses is "
                                      f"{sorted(_CAPABILITIES.wideband_responses)}")

    @property
    def frozen_topology(self) -> PropagationTopology:
        if isinstance(self.topology, PreparedFixedTopology):
            return self.topology.topology
        return self.topology


@dataclass(frozen=True, slots=True, eq=False)
class FixedTopologyEvaluation:
    """A replayed batch of frozen rows (reference consumer.py:1421-1480).
    ``row_valid`` (when present) is the sole authority on whether a row still
    exists; replay is subtractive: paths that appear at the new endpoints are
    not discovered here."""

    paths: PropagationPathBatch
    convention: PropagationConvention
    capabilities: PropagationCapabilities
    diagnostics: PropagationDiagnostics
    row_valid: torch.Tensor | None = None

    def __post_init__(self) -> None:
        if self.row_valid is not None:
            _require_tensor("row_valid", self.row_valid, dtype=torch.bool, shape=(self.paths.path_count,))


def _replay_rows(dscene: Any, topology: PropagationTopology, source: torch.Tensor, target: torch.Tensor,
                 tx_power: torch.Tensor, tx_pol: torch.Tensor, rx_pol: torch.Tensor,
                 frequency_hz: float) -> dict[str, torch.Tensor]:
    """One native launch of wc_fixed_rows_evaluate over every frozen row."""
    k = topology.row_count
    width = int(topology.primitive_sequence.shape[1])
    device = source.device
    out = {
        "valid": torch.empty(k, dtype=torch.uint8, device=device),
        "path_field": torch.empty(k, dtype=torch.complex64, device=device),
        "field_xyz": torch.empty((k, 3), dtype=torch.complex64, device=device),
        "path_length": torch.empty(k, dtype=torch.float32, device=device),
        "delay": torch.empty(k, dtype=torch.float32, device=device),
        "direction": torch.empty((k, 3), dtype=torch.float32, device=device),
        "positions": torch.zeros((k, width, 3), dtype=torch.float32, device=device),
        "normals": torch.zeros((k, width, 3), dtype=torch.float32, device=device),
    }
    if k == 0:
        return out
    p = _native.ptr
    component = topology.component_id.to(torch.int32).contiguous()
    depth = topology.depth.to(torch.int32).contiguous()
    faces = topology.primitive_sequence.to(torch.int32).contiguous()
    view = _native.WcFixedRows(
        count=k, width=width, component=p(component), depth=p(depth), faces=p(faces) if width else None,
        source=p(source), target=p(target), tx_power=p(tx_power), tx_polarization=p(tx_pol),
        rx_polarization=p(rx_pol), valid=p(out["valid"]), path_field=p(torch.view_as_real(out["path_field"])),
        field_xyz=p(torch.view_as_real(out["field_xyz"])), path_length=p(out["path_length"]), delay=p(out["delay"]),
        direction=p(out["direction"]), positions=p(out["positions"]) if width else None,
        normals=p(out["normals"]) if width else None)
    _native.call("wc_fixed_rows_evaluate", dscene.ref(), ctypes.byref(view), float(frequency_hz),
                 _native.stream_handle(device))
    return out


def _image_hits(source: torch.Tensor, target: torch.Tensor, plane: torch.Tensor) -> torch.Tensor:
    """Interaction points of reflection rows over fixed planes, differentiable
    in the endpoints (float64): the source mirrored through each plane in turn,
    then the back-projection from the target through the images. ``plane`` is
    [n, depth, 4] (unit normal, offset); returns [n, depth, 3]."""
    normal, offset = plane[..., :3], plane[..., 3]
    depth = int(plane.shape[1])
    images = []
    image = source
    for j in range(depth):
        dist = (image * normal[:, j]).sum(-1) - offset[:, j]
        image = image - 2.0 * dist[:, None] * normal[:, j]
        images.append(image)
    hits: list[torch.Tensor] = [source] * depth
    nxt = target
    for j in reversed(range(depth)):
        si = (images[j] * normal[:, j]).sum(-1) - offset[:, j]
        sn = (nxt * normal[:, j]).sum(-1) - offset[:, j]
        hits[j] = images[j] + (nxt - images[j]) * (si / (si - sn))[:, None]
        nxt = hits[j]
    return torch.stack(hits, dim=1)


def _rows_ad(out: dict[str, torch.Tensor], valid: torch.Tensor, topo: PropagationTopology, compiled: CompiledScene,
             source: torch.Tensor, target: torch.Tensor, tx_power: torch.Tensor, tx_pol: torch.Tensor,
             rx_pol: torch.Tensor, frequency: Any, frequency_value: float,
             offset: float = 0.0) -> dict[str, torch.Tensor]:
    """Frozen rows through the differentiable row operators (reference
    consumer.py:3994-4017 and the prepared route): line-of-sight rows through
    ``field_free_space_ad``, reflection rows, grouped by depth, through
    ``field_reflection_sequence_ad`` at interaction points re-solved by the
    image method over their frozen faces' planes (``_image_hits``).
    Derivatives reach the endpoint positions and, when the reference frequency
    is a tensor, the frequency (reverse or forward mode); power,
    polarizations, materials and planes enter as constants. Rows the native
    replay found invalid keep zero transport."""
    from witwin.channel.kernels import fields as field_kernels

    device = source.device
    k = topo.row_count
    f = frequency + offset if isinstance(frequency, torch.Tensor) else frequency_value + offset
    fv = frequency_value + offset
    power, p_tx, p_rx = tx_power.detach(), tx_pol.detach(), rx_pol.detach()
    component = topo.component_id.to(torch.int64)
    depth = topo.depth.to(torch.int64)
    pieces = []  # (row indices, operator outputs)
    los = torch.nonzero(component == 0).flatten()
    if los.numel():
        pieces.append((los, field_kernels.field_free_space_ad(source[los], target[los], power[los], p_tx[los],
                                                              p_rx[los], frequency=f, frequency_value=fv)))
    refl = component == 1
    if bool(refl.any()):
        mats = compiled.materials
        table = {name: torch.as_tensor(np.asarray(v, np.float32), device=device) for name, v in
                 (("eps", mats.eps_r), ("sigma", mats.sigma_e), ("mu", mats.mu_r), ("gain", mats.gain),
                  ("thickness", mats.thickness_m))}
        face_group = torch.as_tensor(np.asarray(compiled.face_group, np.int64), device=device)
        face_material = torch.as_tensor(np.asarray(compiled.face_material, np.int64), device=device)
        group_plane = torch.as_tensor(np.asarray(compiled.group_plane, np.float64), device=device)
        for d in sorted(set(depth[refl].tolist())):
            rows = torch.nonzero(refl & (depth == d)).flatten()
            faces = topo.primitive_sequence[rows, :d].to(torch.int64)
            plane = group_plane[face_group[faces]]
            hits = _image_hits(source[rows].double(), target[rows].double(), plane).float()
            m = face_material[faces]
            pieces.append((rows, field_kernels.field_reflection_sequence_ad(
                source[rows], target[rows], hits.contiguous(), plane[..., :3].float().contiguous(), power[rows],
                p_tx[rows], p_rx[rows], table["eps"][m], table["sigma"][m], table["mu"][m], table["gain"][m],
                table["thickness"][m], frequency=f, frequency_value=fv)))

    def assemble(key: str, like: torch.Tensor) -> torch.Tensor:
        full = torch.zeros_like(like)
        for rows, values in pieces:
            full = full.index_copy(0, rows, values[key].to(like.dtype))
        return full

    path_field = assemble("path_field", out["path_field"])
    field_xyz = assemble("field_vector", out["field_xyz"])
    zero = torch.zeros((), dtype=path_field.dtype, device=device)
    return {**out, "path_field": torch.where(valid, path_field, zero),
            "field_xyz": torch.where(valid.reshape(-1, 1), field_xyz, zero),
            "path_length": assemble("path_length_m", out["path_length"]), "delay": assemble("delay_s", out["delay"]),
            "direction": assemble("direction", out["direction"])} if k else out


def reevaluate(compiled_scene: CompiledScene, request: FixedTopologyRequest) -> FixedTopologyEvaluation:
    """Replay frozen rows at the request's endpoints without discovery or
    compaction (reference consumer.py:3965-4110): line-of-sight rows re-test
    their segment, reflection rows re-solve the image method over their frozen
    faces' planes and re-check faces and legs; rows that no longer exist are
    published with ``row_valid = False`` and zero transport."""
    if not isinstance(compiled_scene, CompiledScene):
        raise TypeError("reevaluate requires a CompiledScene")
    if not isinstance(request, FixedTopologyRequest):
        raise TypeError("request must be a FixedTopologyRequest")
    compiled = compiled_scene
    freq = float(request.reference_frequency_hz.detach()) if isinstance(request.reference_frequency_hz,
                                                                         torch.Tensor) \
        else float(request.reference_frequency_hz)
    if abs(freq - compiled.frequency_hz) > 1e-6 * compiled.frequency_hz:
        raise ValueError(f"reference_frequency_hz {freq} does not match the compiled scene's "
                         f"{compiled.frequency_hz}")
    if request.ad_mode not in _CAPABILITIES.fixed_topology_ad_modes:
        raise NotImplementedError(f"ad_mode {request.ad_mode!r} is not supported by reevaluate; supported: "
                                  f"{sorted(_CAPABILITIES.fixed_topology_ad_modes)}")
    topo = request.frozen_topology
    width = int(topo.primitive_sequence.shape[1])
    prepared = request.topology if isinstance(request.topology, PreparedFixedTopology) else None
    if prepared is None and width != 0:
        raise NotImplementedError("fixed LoS reevaluation requires zero-width interaction sequences; call "
                                  "prepare_fixed_topology first to reevaluate a topology that carries interactions")
    if prepared is not None and request.slot_count != prepared.slot_count:
        raise ValueError(f"slot_count={request.slot_count} does not match the topology's "
                         f"{prepared.slot_count} slots (replicate_over_slots)")
    components = {b.component for b in prepared.buckets} if prepared is not None else {"los"}
    if request.response == "polarimetric_transport":
        if request.sources.polarization_basis is None or request.sinks.polarization_basis is None:
            raise ValueError("polarimetric_transport requires source and sink polarization_basis tensors")
        if components - {"los"}:
            raise NotImplementedError("polarimetric_transport replays line-of-sight rows only")
    differentiable = request.ad_mode != "none"
    if differentiable:
        if request.response == "polarimetric_transport":
            raise NotImplementedError(f"ad_mode {request.ad_mode!r} covers the scalar and complex3 transports")
        if components - _CAPABILITIES.fixed_topology_ad_components:
            raise NotImplementedError(f"ad_mode {request.ad_mode!r} replays line-of-sight and reflection rows; the rows carry "
                                      f"{sorted(components - _CAPABILITIES.fixed_topology_ad_components)}")
    if topo.primitive_sequence.device != request.sources.positions_m.device:
        raise ValueError("fixed topology and endpoint batches must share a device")
    device = _native.require_gpu()
    dscene = device_scene(compiled, device)
    sources, sinks = request.sources, request.sinks
    k = topo.row_count
    slots = request.slot_count
    s_count = sources.count // slots
    r_count = sinks.count // slots
    src = topo.source_index.to(torch.int64)
    snk = topo.sink_index.to(torch.int64)
    # the one validation read: every index in range and, with slots, inside its row's slot
    if k:
        slot_of_row = torch.arange(k, device=src.device) // max(k // slots, 1)
        bad = (src < 0) | (src >= sources.count) | (snk < 0) | (snk >= sinks.count)
        if slots > 1:
            bad = bad | (src // s_count != slot_of_row) | (snk // r_count != slot_of_row)
        if bool(bad.any().item()):
            raise ValueError("frozen rows name endpoints outside the request's batches (or outside their slot)")
    source = sources.positions_m[src].contiguous()
    target = sinks.positions_m[snk].contiguous()
    tx_power = sources.powers_w[src].to(torch.float32).contiguous()
    tx_pol = sources.polarizations[src].to(torch.float32).contiguous()
    rx_pol = sinks.polarizations[snk].to(torch.float32).contiguous()
    out = _replay_rows(dscene, topo, source, target, tx_power, tx_pol, rx_pol, compiled.frequency_hz)
    valid = out["valid"].bool()
    if differentiable:
        # the native replay decided which rows still see each other; the
        # values of those rows come from the differentiable free-space operator
        out = _rows_ad(out, valid, topo, compiled, source, target, tx_power, tx_pol, rx_pol,
                       request.reference_frequency_hz, compiled.frequency_hz)
    amplitude = torch.sqrt(torch.clamp(tx_power, min=0.0)).unsqueeze(-1)
    offsets = request.frequency_offsets_hz
    if slots > 1:
        slot = src // s_count
        pair_index = slot * (s_count * r_count) + (snk - slot * r_count) * s_count + (src - slot * s_count)
        pair_count = slots * s_count * r_count
    else:
        pair_index = snk * sources.count + src
        pair_count = sources.count * sinks.count
    counts = torch.bincount(pair_index, minlength=pair_count) if pair_count else \
        torch.zeros(0, dtype=torch.int64, device=device)
    pair_offsets = torch.zeros(pair_count + 1, dtype=torch.int64, device=device)
    if pair_count:
        pair_offsets[1:] = torch.cumsum(counts, 0)
    geometry = PropagationGeometry(path_length_m=out["path_length"], delay_s=out["delay"],
                                   field_direction=out["direction"], interaction_positions_m=out["positions"],
                                   interaction_normals=out["normals"])
    columns = []
    if offsets is not None:
        for df in offsets:
            f = compiled.frequency_hz + df
            if not f > 0.0:
                raise ValueError(f"frequency offset {df} gives a non-positive frequency")
            column = _replay_rows(dscene, topo, source, target, tx_power, tx_pol, rx_pol, f)
            if differentiable:
                column = _rows_ad(column, valid, topo, compiled, source, target, tx_power, tx_pol, rx_pol,
                                  request.reference_frequency_hz, compiled.frequency_hz, df)
            columns.append(column)
    if request.response == "scalar_transport":
        payload = torch.stack([c["path_field"] for c in columns], dim=1) if columns else None
        transport: Any = ScalarTransport(coefficient=out["path_field"], coefficient_offsets=payload,
                                         frequency_offsets_hz=offsets)
    elif request.response == "complex3_transport":
        payload = torch.stack([c["field_xyz"] * amplitude for c in columns], dim=1) if columns else None
        transport = Complex3Transport(field=(out["field_xyz"] * amplitude).contiguous(), direction=out["direction"],
                                      field_offsets=payload, frequency_offsets_hz=offsets)
    else:
        transport = _los_jones(source, target, sources.polarization_basis[src], sinks.polarization_basis[snk],
                               compiled.frequency_hz)
        zero = torch.zeros((), dtype=transport.matrix.dtype, device=device)
        transport = JonesTransport(matrix=torch.where(valid.reshape(-1, 1, 1), transport.matrix, zero),
                                   source_basis=transport.source_basis, sink_basis=transport.sink_basis)
    diagnostics = PropagationDiagnostics(
        discovery_launch_count=0, candidate_count=0, visibility_rejection_count=0, compact_count_d2h_copies=0,
        compact_count_d2h_bytes=0, compact_sync_count=0, validation_d2h_copies=1 if k else 0,
        validation_d2h_bytes=1 if k else 0, validation_sync_count=1 if k else 0,
        frequency_column_count=1 if offsets is None else len(offsets))
    paths = PropagationPathBatch(pair_count=pair_count, path_count=k, pair_index=pair_index.contiguous(),
                                 pair_offsets=pair_offsets, topology=topo, geometry=geometry, transport=transport)
    publish = bool(components & _CAPABILITIES.fixed_topology_row_validity_components)
    return FixedTopologyEvaluation(paths=paths, convention=PropagationConvention(), capabilities=_CAPABILITIES,
                                   diagnostics=diagnostics, row_valid=valid if publish else None)


__all__ = [
    "AD_MODES", "COMPONENTS", "CONTRACT_VERSION", "Complex3Transport", "EndpointBatch", "JonesTransport",
    "MAX_DEPTH", "PropagationCapabilities", "PropagationConvention", "PropagationDiagnostics",
    "PropagationEvaluation", "PropagationGeometry", "PropagationPathBatch", "PropagationRequest",
    "PropagationTopology", "RESPONSES", "ScalarTransport", "TOPOLOGY_MODES", "capabilities", "evaluate",
    "transverse_basis", "WORLD_MOTIONS", "FixedTopologyBucket", "PreparedFixedTopology", "prepare_fixed_topology",
    "replicate_over_slots", "FixedTopologyRequest", "FixedTopologyEvaluation", "reevaluate",
]
