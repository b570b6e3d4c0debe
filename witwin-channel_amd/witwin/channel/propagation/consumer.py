"""Solver-neutral propagation consumer: explicit source and sink endpoint
batches against a compiled scene, compact path rows with exact pair
segmentation, and the scalar / complex3 / polarimetric (Jones) transports.

Mirrors the reference consumer contract (witwin/channel/propagation/
consumer.py): the vocabulary (:405-435), ``EndpointBatch`` (:515-588),
``PropagationRequest`` (:623-653), ``PropagationTopology`` /
``PropagationGeometry`` (:656-793), the transports (:796-916),
``PropagationPathBatch`` (:919-960), ``PropagationConvention`` (:963-1005),
``capabilities`` (:1181) and ``evaluate`` (:3478-3520).

The rows come from the same HIP engine as ``deterministic.solve`` with
``export_paths`` (count -> one count read -> emit -> canonical order), run on
the request's endpoints instead of the compiled scene's: rows are pair-major
with ``pair_index = sink_index * source_count + source_index``. The Jones
response is line-of-sight only, as in the reference; it excites the native
free-space row operator with the two source basis vectors and projects onto
the sink basis. Forward evaluation only: ``ad_mode`` other than ``"none"`` is
refused (the differentiable route of this build is ``deterministic.solve``
with ``ad_mode="vjp"``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any

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
    supports_fixed_topology: bool = False

    def components_for(self, response: str) -> frozenset[str]:
        return dict(self.response_components)[response]

    def ad_modes_for(self, response: str) -> frozenset[str]:
        return dict(self.response_ad_modes)[response]


_CAPABILITIES = PropagationCapabilities(
    contract_version=CONTRACT_VERSION,
    components=COMPONENTS,
    responses=RESPONSES,
    topology_modes=TOPOLOGY_MODES,
    ad_modes=frozenset({"none"}),
    response_components=(("scalar_transport", COMPONENTS), ("complex3_transport", COMPONENTS),
                         ("polarimetric_transport", frozenset({"los"}))),
    response_ad_modes=(("scalar_transport", frozenset({"none"})), ("complex3_transport", frozenset({"none"})),
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
                     ctypes.c_float(frequency_hz), p(torch.view_as_real(out["field_vector"])),
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


__all__ = [
    "AD_MODES", "COMPONENTS", "CONTRACT_VERSION", "Complex3Transport", "EndpointBatch", "JonesTransport",
    "MAX_DEPTH", "PropagationCapabilities", "PropagationConvention", "PropagationDiagnostics",
    "PropagationEvaluation", "PropagationGeometry", "PropagationPathBatch", "PropagationRequest",
    "PropagationTopology", "RESPONSES", "ScalarTransport", "TOPOLOGY_MODES", "capabilities", "evaluate",
    "transverse_basis",
]
