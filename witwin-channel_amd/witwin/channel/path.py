"""Explicit path solve: per-(receiver, transmitter) path sets packed into a
padded ``PathResult`` with CIR / CFR / tap views.

Mirrors the reference's public path API (witwin/channel/path.py): ``Config``
(:76-151), ``RaggedPathSoA`` (:164-329), ``InteractionType`` (:332-337),
``PathResult`` (:354-788) with ``cir`` (:637), ``cfr`` (:646), ``taps``
(:658), ``beamform`` (:701), ``filter_by_type`` (:734), ``endpoint_angles``
(:879), ``from_evaluated_paths`` (:887-975) and ``solve`` (:1707-1719).

The paths come from the same HIP export as the deterministic solver
(count -> emit -> canonical order; one host read of the row count), so the
topology, the canonical row order and the pair segmentation are identical to
``deterministic.solve(..., export_paths=True)``. ``PathResult.cfr`` runs the
``wc_paths_cfr`` kernel (phase reduced in float64). Endpoints are single
isotropic elements (num_rx_ant = num_tx_ant = 1); antenna arrays are outside
this build.
"""

from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, replace
from typing import Any

import numpy as np
import torch

from witwin.channel import _native
from witwin.channel.components import (
    AD_MODES,
    DEFAULT_COMPONENTS,
    DEPTH_CAPPED_COMPONENTS,
    component_bits,
    component_max_depth,
    validate_max_depth,
    validated_components,
)
from witwin.channel.constants import PHASE_CONVENTION
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import CompiledScene, compile as compile_scene, device_scene

LIGHT_SPEED_M_PER_S = 299_792_458.0
_MAX_PATHS_SCOPES = frozenset({"per_pair"})


@dataclass(frozen=True, slots=True)
class Config:
    """Path-solve configuration (reference path.py:76-151)."""

    max_depth: int = 1
    components: frozenset[str] | set[str] | tuple[str, ...] | list[str] = DEFAULT_COMPONENTS
    max_paths: int | None = None
    max_paths_scope: str = "per_pair"
    ad_mode: str = "none"
    coupled_paths: bool = False
    coupled_candidate_limit: int = 1_000_000
    scattering_samples_per_m2: float = 8.0
    scattering_max_paths_per_pair: int = 4096
    scattering_power_threshold: float = 0.0
    scattering_chain_max_depth: int = 0
    scattering_chain_samples_per_m2: float = 2.0
    scattering_chain_max_rows: int = 256
    isb_boundary_taper: bool = False
    isb_boundary_taper_width: float = 0.5

    def __post_init__(self) -> None:
        validate_max_depth(self.max_depth)
        if not (0.0 < self.isb_boundary_taper_width <= 4.0):
            raise ValueError("isb_boundary_taper_width must be in (0, 4]")
        if self.scattering_samples_per_m2 <= 0.0:
            raise ValueError("scattering_samples_per_m2 must be positive")
        if self.scattering_max_paths_per_pair <= 0:
            raise ValueError("scattering_max_paths_per_pair must be positive")
        if self.scattering_power_threshold < 0.0:
            raise ValueError("scattering_power_threshold must be non-negative")
        components = validated_components(self.components,
                                          error_message="components must be a subset of {valid}")
        if self.max_depth > 5 and components & DEPTH_CAPPED_COMPONENTS:
            raise RuntimeError("path reflection/transmission support max_depth <= 5")
        if self.coupled_paths:
            if self.max_depth < 2:
                raise RuntimeError("coupled reflection-diffraction paths require max_depth >= 2")
            if not {"reflection", "diffraction"}.issubset(components):
                raise RuntimeError("coupled paths require both reflection and diffraction components")
        if self.max_paths is not None and self.max_paths <= 0:
            raise ValueError("max_paths must be positive when set")
        if self.max_paths_scope not in _MAX_PATHS_SCOPES:
            raise ValueError("path max_paths_scope must be 'per_pair'")
        if self.coupled_candidate_limit <= 0:
            raise ValueError("coupled_candidate_limit must be positive")
        if self.ad_mode not in AD_MODES:
            raise ValueError(f"path ad_mode must be one of {sorted(AD_MODES)}")
        object.__setattr__(self, "components", components)


class InteractionType(enum.IntFlag):
    NONE = 0
    REFLECTION = 1
    DIFFRACTION = 2
    TRANSMISSION = 4
    SCATTERING = 8


def _check(predicate: torch.Tensor, message: str) -> None:
    if predicate.device.type == "cuda":
        torch._assert_async(predicate, message)
    elif not bool(predicate):
        raise ValueError(message)


@dataclass(frozen=True, slots=True)
class RaggedPathSoA:
    """Per-link ragged path storage in pair order (reference path.py:164-329).
    Pair index = ((rx * num_rx_ant + rx_ant) * num_tx + tx) * num_tx_ant + tx_ant;
    paths keep their input order inside a pair."""

    num_rx: int
    num_rx_ant: int
    num_tx: int
    num_tx_ant: int
    num_time_steps: int
    pair_offsets: torch.Tensor
    rx_id: torch.Tensor
    rx_ant_id: torch.Tensor
    tx_id: torch.Tensor
    tx_ant_id: torch.Tensor
    field: torch.Tensor
    delay_s: torch.Tensor
    theta_t: torch.Tensor
    phi_t: torch.Tensor
    theta_r: torch.Tensor
    phi_r: torch.Tensor
    interaction_type: torch.Tensor
    primitive_id: torch.Tensor
    material_id: torch.Tensor
    position: torch.Tensor
    normal: torch.Tensor

    @property
    def path_count(self) -> int:
        return int(self.delay_s.shape[0])

    @property
    def pair_count(self) -> int:
        return self.num_rx * self.num_rx_ant * self.num_tx * self.num_tx_ant

    @property
    def max_depth(self) -> int:
        return int(self.interaction_type.shape[1])

    @classmethod
    def from_flat(cls, *, num_rx: int, num_rx_ant: int, num_tx: int, num_tx_ant: int, rx_id: torch.Tensor,
                  tx_id: torch.Tensor, field: torch.Tensor, delay_s: torch.Tensor, theta_t: torch.Tensor,
                  phi_t: torch.Tensor, theta_r: torch.Tensor, phi_r: torch.Tensor,
                  interaction_type: torch.Tensor, primitive_id: torch.Tensor, material_id: torch.Tensor,
                  position: torch.Tensor, normal: torch.Tensor, rx_ant_id: torch.Tensor | None = None,
                  tx_ant_id: torch.Tensor | None = None,
                  max_paths_per_pair: int | None = None) -> "RaggedPathSoA":
        count = int(delay_s.shape[0])
        device = delay_s.device
        if field.ndim == 1:
            field = field.unsqueeze(-1)
        if field.ndim != 2 or field.shape[0] != count:
            raise ValueError("field must have shape (path, time)")
        if field.dtype != torch.complex64:
            raise ValueError("field must use complex64")
        if max_paths_per_pair is not None and max_paths_per_pair <= 0:
            raise ValueError("max_paths_per_pair must be positive when set")
        zeros = lambda: torch.zeros((count,), device=device, dtype=torch.int32)  # noqa: E731
        rx_ant_id = zeros() if rx_ant_id is None else rx_ant_id.to(device=device, dtype=torch.int32)
        tx_ant_id = zeros() if tx_ant_id is None else tx_ant_id.to(device=device, dtype=torch.int32)
        rx_id = rx_id.to(device=device, dtype=torch.int32)
        tx_id = tx_id.to(device=device, dtype=torch.int32)
        for name, value in (("rx_id", rx_id), ("rx_ant_id", rx_ant_id), ("tx_id", tx_id),
                            ("tx_ant_id", tx_ant_id), ("delay_s", delay_s), ("theta_t", theta_t),
                            ("phi_t", phi_t), ("theta_r", theta_r), ("phi_r", phi_r)):
            if value.shape != (count,):
                raise ValueError(f"{name} must have shape (path,)")
        if interaction_type.ndim != 2:
            raise ValueError("interaction_type must have shape (path, depth)")
        depth = int(interaction_type.shape[1])
        for name, value in (("interaction_type", interaction_type), ("primitive_id", primitive_id),
                            ("material_id", material_id)):
            if value.shape != (count, depth):
                raise ValueError(f"{name} must have shape (path, depth)")
        for name, value in (("position", position), ("normal", normal)):
            if value.shape != (count, depth, 3):
                raise ValueError(f"{name} must have shape (path, depth, 3)")
        pair_count = int(num_rx) * int(num_rx_ant) * int(num_tx) * int(num_tx_ant)
        if count and pair_count <= 0:
            raise ValueError("non-empty paths require non-empty endpoint dimensions")
        for ids, dim, name in ((rx_id, num_rx, "rx_id"), (rx_ant_id, num_rx_ant, "rx_ant_id"),
                               (tx_id, num_tx, "tx_id"), (tx_ant_id, num_tx_ant, "tx_ant_id")):
            _check(((ids >= 0) & (ids < int(dim))).all(), f"{name} is outside the declared endpoint dimension")
        pair = ((rx_id.long() * int(num_rx_ant) + rx_ant_id) * int(num_tx) + tx_id) * int(num_tx_ant) + tx_ant_id
        order = torch.argsort(pair, stable=True)
        pair = pair[order]
        counts = torch.bincount(pair, minlength=pair_count)
        if max_paths_per_pair is not None:
            starts = torch.cumsum(counts, 0) - counts
            rank = torch.arange(count, device=device) - torch.repeat_interleave(starts, counts)
            keep = rank < int(max_paths_per_pair)
            order, pair = order[keep], pair[keep]
            counts = torch.bincount(pair, minlength=pair_count)
        offsets = torch.zeros(pair_count + 1, device=device, dtype=torch.int64)
        offsets[1:] = torch.cumsum(counts, 0)

        def take(value: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
            return value.to(device=device, dtype=dtype)[order].contiguous()

        return cls(num_rx=int(num_rx), num_rx_ant=int(num_rx_ant), num_tx=int(num_tx), num_tx_ant=int(num_tx_ant),
                   num_time_steps=int(field.shape[1]), pair_offsets=offsets,
                   rx_id=take(rx_id, torch.int32), rx_ant_id=take(rx_ant_id, torch.int32),
                   tx_id=take(tx_id, torch.int32), tx_ant_id=take(tx_ant_id, torch.int32),
                   field=take(field, torch.complex64), delay_s=take(delay_s, torch.float32),
                   theta_t=take(theta_t, torch.float32), phi_t=take(phi_t, torch.float32),
                   theta_r=take(theta_r, torch.float32), phi_r=take(phi_r, torch.float32),
                   interaction_type=take(interaction_type, torch.int32),
                   primitive_id=take(primitive_id, torch.int32), material_id=take(material_id, torch.int32),
                   position=take(position, torch.float32), normal=take(normal, torch.float32))


def _masked_min(values: torch.Tensor, valid: torch.Tensor) -> torch.Tensor:
    """Per-pair minimum over valid paths (0 for a pair without paths)."""
    if values.shape[-1] == 0:
        return torch.zeros((*values.shape[:-1], 1), device=values.device, dtype=values.dtype)
    low = torch.where(valid, values, torch.full_like(values, float("inf"))).min(dim=-1, keepdim=True).values
    return torch.where(torch.isfinite(low), low, torch.zeros_like(low))


def _uniform_grid(frequencies: torch.Tensor) -> tuple[float, float] | None:
    """(f0, df) when a host-resident frequency list is an arithmetic grid to
    1e-12 of its largest value (a phase difference below 1e-6 rad for delays
    up to 10 us at 28 GHz)."""
    if frequencies.device.type != "cpu" or frequencies.numel() < 2:
        return None
    f = frequencies.detach().to(torch.float64).numpy()
    df = (f[-1] - f[0]) / (f.size - 1)
    grid = f[0] + df * np.arange(f.size)
    if df != 0.0 and np.max(np.abs(grid - f)) <= 1e-12 * np.max(np.abs(f)):
        return float(f[0]), float(df)
    return None


def paths_cfr(a: torch.Tensor, tau: torch.Tensor, valid: torch.Tensor, frequencies: torch.Tensor, *,
              normalize_delays: bool = True) -> torch.Tensor:
    """``wc_paths_cfr`` over padded path sets: a [B, P, S] c64, tau / valid [B, P]
    -> [B, S, F] c64 (sum over valid paths of a * exp(-j 2 pi f (tau - tau0)))."""

    _native.require_gpu()
    if frequencies.ndim != 1:
        raise ValueError("frequencies must have shape (frequency,)")
    device = a.device
    b, p, s = (int(x) for x in a.shape)
    f_count = int(frequencies.shape[0])
    out = torch.empty((b, s, f_count), device=device, dtype=torch.complex64)
    grid = _uniform_grid(frequencies)
    freq_dev = None if grid is not None else frequencies.detach().to(device=device, dtype=torch.float64).contiguous()
    f0, df = grid if grid is not None else (0.0, 0.0)
    a_c = a.contiguous()
    tau_c = tau.to(torch.float32).contiguous()
    valid_c = valid.to(torch.uint8).contiguous()
    p_ = _native.ptr
    _native.call("wc_paths_cfr", p_(torch.view_as_real(a_c)) if a_c.numel() else None, p_(tau_c), p_(valid_c), b, p,
                 s, p_(freq_dev), ctypes.c_double(f0), ctypes.c_double(df), f_count, 1 if normalize_delays else 0,
                 p_(torch.view_as_real(out)) if out.numel() else None, _native.stream_handle(device))
    return out


@dataclass(frozen=True, slots=True)
class PathResult:
    """Padded public path result, shapes (rx, rx_ant, tx, tx_ant, path[, time])
    (reference path.py:354-788)."""

    a: torch.Tensor
    tau: torch.Tensor
    theta_t: torch.Tensor
    phi_t: torch.Tensor
    theta_r: torch.Tensor
    phi_r: torch.Tensor
    valid: torch.Tensor
    interaction_type: torch.Tensor
    primitive_id: torch.Tensor
    material_id: torch.Tensor
    position: torch.Tensor
    normal: torch.Tensor
    num_paths: torch.Tensor
    metadata: dict[str, Any]
    field_xyz: torch.Tensor | None = None
    field_direction: torch.Tensor | None = None
    tx_weights: torch.Tensor | None = None
    rx_weights: torch.Tensor | None = None

    def __post_init__(self) -> None:
        if self.a.ndim != 6 or self.a.dtype != torch.complex64:
            raise ValueError("a must be complex64 with shape (rx, rx_ant, tx, tx_ant, path, time)")
        path_shape = tuple(self.a.shape[:-1])
        for name in ("tau", "theta_t", "phi_t", "theta_r", "phi_r", "valid"):
            if tuple(getattr(self, name).shape) != path_shape:
                raise ValueError(f"{name} must match the path shape")
        depth_shape = (*path_shape, self.max_depth)
        for name in ("interaction_type", "primitive_id", "material_id"):
            if tuple(getattr(self, name).shape) != depth_shape:
                raise ValueError(f"{name} must match the depth shape")
        for name in ("position", "normal"):
            if tuple(getattr(self, name).shape) != (*depth_shape, 3):
                raise ValueError(f"{name} must match the vector depth shape")
        if tuple(self.num_paths.shape) != self.path_count_shape:
            raise ValueError("num_paths must match the endpoint-pair shape")
        if self.field_xyz is None:
            object.__setattr__(self, "field_xyz", torch.zeros((*path_shape, 3), device=self.a.device,
                                                              dtype=torch.complex64))
        if self.field_direction is None:
            object.__setattr__(self, "field_direction", torch.zeros((*path_shape, 3), device=self.a.device,
                                                                    dtype=torch.float32))
        if tuple(self.field_xyz.shape) != (*path_shape, 3) or self.field_xyz.dtype != torch.complex64:
            raise ValueError("field_xyz must be complex64 matching the path shape with a complex3 tail")
        if tuple(self.field_direction.shape) != (*path_shape, 3) or self.field_direction.dtype != torch.float32:
            raise ValueError("field_direction must be float32 matching the path shape with a vec3 tail")
        if self.valid.dtype != torch.bool:
            raise ValueError("valid must use bool")
        if any(getattr(self, n).dtype != torch.float32 for n in ("tau", "theta_t", "phi_t", "theta_r", "phi_r",
                                                                  "position", "normal")):
            raise ValueError("path geometry tensors must use float32")
        if any(getattr(self, n).dtype != torch.int32 for n in ("interaction_type", "primitive_id", "material_id",
                                                                "num_paths")):
            raise ValueError("interaction and path-count tensors must use int32")

    @property
    def num_rx(self) -> int:
        return int(self.a.shape[0])

    @property
    def num_rx_ant(self) -> int:
        return int(self.a.shape[1])

    @property
    def num_tx(self) -> int:
        return int(self.a.shape[2])

    @property
    def num_tx_ant(self) -> int:
        return int(self.a.shape[3])

    @property
    def max_num_paths(self) -> int:
        return int(self.a.shape[4])

    @property
    def num_time_steps(self) -> int:
        return int(self.a.shape[5])

    @property
    def max_depth(self) -> int:
        return int(self.interaction_type.shape[-1])

    @property
    def path_shape(self) -> tuple[int, int, int, int, int]:
        return tuple(int(v) for v in self.a.shape[:-1])  # type: ignore[return-value]

    @property
    def path_count_shape(self) -> tuple[int, int, int, int]:
        return tuple(int(v) for v in self.a.shape[:4])  # type: ignore[return-value]

    @property
    def rx_id(self) -> torch.Tensor:
        ids = torch.arange(self.num_rx, device=self.a.device, dtype=torch.int32)
        return ids.reshape(-1, 1, 1, 1, 1).expand(self.path_shape).contiguous()

    @property
    def tx_id(self) -> torch.Tensor:
        ids = torch.arange(self.num_tx, device=self.a.device, dtype=torch.int32)
        return ids.reshape(1, 1, -1, 1, 1).expand(self.path_shape).contiguous()

    @property
    def path_length_m(self) -> torch.Tensor:
        length = self.tau * LIGHT_SPEED_M_PER_S
        return torch.where(self.valid, length, torch.full_like(length, -1.0))

    @property
    def types(self) -> torch.Tensor:
        return self.interaction_type

    @property
    def vertices(self) -> torch.Tensor:
        return self.position

    @property
    def normals(self) -> torch.Tensor:
        return self.normal

    @property
    def objects(self) -> torch.Tensor:
        return self.primitive_id

    @classmethod
    def from_ragged(cls, ragged: RaggedPathSoA, *, max_paths_per_pair: int | None = None,
                    minimum_path_width: int = 0, metadata: dict[str, Any] | None = None) -> "PathResult":
        counts = ragged.pair_offsets[1:] - ragged.pair_offsets[:-1]
        if max_paths_per_pair is None:
            # the padded width is data dependent: one host read
            max_paths_per_pair = int(counts.max().item()) if counts.numel() else 0
        width = max(int(max_paths_per_pair), int(minimum_path_width))
        if width < 0:
            raise ValueError("max_paths_per_pair must be non-negative")
        _check((counts <= width).all(), "ragged input exceeds max_paths_per_pair")
        base = (ragged.num_rx, ragged.num_rx_ant, ragged.num_tx, ragged.num_tx_ant)
        shape = (*base, width)
        device = ragged.delay_s.device
        rank = torch.arange(ragged.path_count, device=device) - torch.repeat_interleave(ragged.pair_offsets[:-1],
                                                                                        counts)
        pair = torch.repeat_interleave(torch.arange(ragged.pair_count, device=device), counts)
        slot = pair * width + rank
        total = ragged.pair_count * width

        def pad(source: torch.Tensor, tail: tuple[int, ...], fill: float | int) -> torch.Tensor:
            out = torch.full((total, *tail), fill, device=device, dtype=source.dtype)
            out[slot] = source
            return out.reshape(*shape, *tail)

        valid = torch.zeros(total, device=device, dtype=torch.bool)
        valid[slot] = True
        meta = dict(metadata or {})
        meta.update({"schema": "PathResult", "schema_version": 1, "max_paths_scope": "per_pair",
                     "stable_order": "input_order_within_receiver_antenna_transmitter_antenna_pair"})
        d = ragged.max_depth
        return cls(a=pad(ragged.field, (ragged.num_time_steps,), 0.0), tau=pad(ragged.delay_s, (), -1.0),
                   theta_t=pad(ragged.theta_t, (), 0.0), phi_t=pad(ragged.phi_t, (), 0.0),
                   theta_r=pad(ragged.theta_r, (), 0.0), phi_r=pad(ragged.phi_r, (), 0.0),
                   valid=valid.reshape(shape),
                   interaction_type=pad(ragged.interaction_type, (d,), int(InteractionType.NONE)),
                   primitive_id=pad(ragged.primitive_id, (d,), -1), material_id=pad(ragged.material_id, (d,), -1),
                   position=pad(ragged.position, (d, 3), 0.0), normal=pad(ragged.normal, (d, 3), 0.0),
                   num_paths=counts.to(torch.int32).reshape(base), metadata=meta)

    def _delays(self, normalize_delays: bool) -> torch.Tensor:
        tau = self.tau
        if normalize_delays:
            tau = tau - _masked_min(tau, self.valid)
        return tau

    def cir(self, *, normalize_delays: bool = True) -> tuple[torch.Tensor, torch.Tensor]:
        tau = torch.where(self.valid, self._delays(normalize_delays), torch.full_like(self.tau, -1.0))
        return torch.where(self.valid.unsqueeze(-1), self.a, torch.zeros_like(self.a)), tau

    def cfr(self, frequencies: torch.Tensor, *, normalize_delays: bool = True) -> torch.Tensor:
        """H[rx, rx_ant, tx, tx_ant, time, f] = sum over valid paths of
        a * exp(-j 2 pi f tau) (delays relative to each pair's first arrival
        when ``normalize_delays``), on the ``wc_paths_cfr`` kernel."""
        if frequencies.ndim != 1:
            raise ValueError("frequencies must have shape (frequency,)")
        pairs = self.num_rx * self.num_rx_ant * self.num_tx * self.num_tx_ant
        out = paths_cfr(self.a.reshape(pairs, self.max_num_paths, self.num_time_steps),
                        self.tau.reshape(pairs, self.max_num_paths), self.valid.reshape(pairs, self.max_num_paths),
                        frequencies, normalize_delays=normalize_delays)
        return out.reshape(*self.path_count_shape, self.num_time_steps, int(frequencies.shape[0]))

    def taps(self, bandwidth: float, num_taps: int, *, normalize_delays: bool = True) -> torch.Tensor:
        if bandwidth <= 0.0:
            raise ValueError("bandwidth must be > 0")
        if num_taps <= 0:
            raise ValueError("num_taps must be > 0")
        tap = torch.round(self._delays(normalize_delays) * float(bandwidth)).to(torch.int64)
        base = self.num_rx * self.num_rx_ant * self.num_tx * self.num_tx_ant
        out = torch.zeros((base, self.num_time_steps, int(num_taps)), device=self.a.device, dtype=torch.complex64)
        if self.max_num_paths == 0:
            return out.reshape(*self.path_count_shape, self.num_time_steps, int(num_taps))
        coeff = self.a.reshape(base, self.max_num_paths, self.num_time_steps)
        tap = tap.reshape(base, self.max_num_paths)
        keep = self.valid.reshape(base, self.max_num_paths) & (tap >= 0) & (tap < num_taps)
        b_idx, p_idx = torch.nonzero(keep, as_tuple=True)
        t_idx = tap[b_idx, p_idx]
        for s in range(self.num_time_steps):
            out[:, s, :].index_put_((b_idx, t_idx), coeff[b_idx, p_idx, s], accumulate=True)
        return out.reshape(*self.path_count_shape, self.num_time_steps, int(num_taps))

    def beamform(self, *, tx_weights: torch.Tensor | None = None,
                 rx_weights: torch.Tensor | None = None) -> "BeamformedPathResult":
        tx = self.tx_weights if tx_weights is None else tx_weights
        rx = self.rx_weights if rx_weights is None else rx_weights
        if tx is None or rx is None:
            raise ValueError("beamform requires tx_weights and rx_weights, either explicitly "
                             "or on every solved endpoint")
        tx = tx.to(device=self.a.device, dtype=torch.complex64)
        rx = rx.to(device=self.a.device, dtype=torch.complex64)
        if tx.ndim == 1 and self.num_tx == 1:
            tx = tx.unsqueeze(0)
        if rx.ndim == 1 and self.num_rx == 1:
            rx = rx.unsqueeze(0)
        if tuple(tx.shape) != (self.num_tx, self.num_tx_ant):
            raise ValueError(f"tx_weights must have shape {(self.num_tx, self.num_tx_ant)}")
        if tuple(rx.shape) != (self.num_rx, self.num_rx_ant):
            raise ValueError(f"rx_weights must have shape {(self.num_rx, self.num_rx_ant)}")
        return BeamformedPathResult(source=self, tx_weights=tx, rx_weights=rx)

    def filter_by_type(self, *interaction_types: int) -> "PathResult":
        """Keep paths containing any of the requested interaction types (0 keeps
        paths without interactions, i.e. line of sight)."""
        if not interaction_types:
            return self
        requested = {int(v) for v in interaction_types}
        has_any = (self.interaction_type != int(InteractionType.NONE)).any(dim=-1)
        keep = self.valid & ~has_any if 0 in requested else torch.zeros_like(self.valid)
        for value in requested - {0}:
            keep = keep | (self.valid & (self.interaction_type == value).any(dim=-1))
        per_rx = self.num_rx_ant * self.num_tx * self.num_tx_ant
        pair = torch.arange(self.num_rx * per_rx, device=self.a.device).view(-1, 1).expand(-1, self.max_num_paths)
        pair = pair.reshape(-1)[keep.reshape(-1)]
        rx_id, rest = pair // per_rx, pair % per_rx
        rx_ant_id, rest = rest // (self.num_tx * self.num_tx_ant), rest % (self.num_tx * self.num_tx_ant)
        tx_id, tx_ant_id = rest // self.num_tx_ant, rest % self.num_tx_ant
        ragged = RaggedPathSoA.from_flat(
            num_rx=self.num_rx, num_rx_ant=self.num_rx_ant, num_tx=self.num_tx, num_tx_ant=self.num_tx_ant,
            rx_id=rx_id, rx_ant_id=rx_ant_id, tx_id=tx_id, tx_ant_id=tx_ant_id, field=self.a[keep],
            delay_s=self.tau[keep], theta_t=self.theta_t[keep], phi_t=self.phi_t[keep],
            theta_r=self.theta_r[keep], phi_r=self.phi_r[keep], interaction_type=self.interaction_type[keep],
            primitive_id=self.primitive_id[keep], material_id=self.material_id[keep],
            position=self.position[keep], normal=self.normal[keep])
        meta = dict(self.metadata)
        meta["filtered_interaction_types"] = sorted(requested)
        filtered = PathResult.from_ragged(ragged, minimum_path_width=1, metadata=meta)
        return replace(filtered, tx_weights=self.tx_weights, rx_weights=self.rx_weights)


@dataclass(frozen=True, slots=True)
class BeamformedPathResult:
    """Beamformed views of a raw ``PathResult`` (reference path.py:791-876)."""

    source: PathResult
    tx_weights: torch.Tensor
    rx_weights: torch.Tensor

    def cir(self, *, normalize_delays: bool = True) -> tuple[torch.Tensor, torch.Tensor]:
        src = self.source
        factor = (self.rx_weights.conj().reshape(src.num_rx, src.num_rx_ant, 1, 1, 1, 1)
                  * self.tx_weights.reshape(1, 1, src.num_tx, src.num_tx_ant, 1, 1))
        coeff = torch.where(src.valid.unsqueeze(-1), src.a * factor, torch.zeros_like(src.a))
        coeff = coeff.permute(0, 2, 1, 3, 4, 5).reshape(src.num_rx, src.num_tx,
                                                        src.num_rx_ant * src.num_tx_ant * src.max_num_paths,
                                                        src.num_time_steps)
        valid = src.valid.permute(0, 2, 1, 3, 4).reshape(src.num_rx, src.num_tx, -1)
        tau = src.tau.permute(0, 2, 1, 3, 4).reshape(src.num_rx, src.num_tx, -1)
        if normalize_delays:
            tau = tau - _masked_min(tau, valid)
        return coeff, torch.where(valid, tau, torch.full_like(tau, -1.0))

    def cfr(self, frequencies: torch.Tensor, *, normalize_delays: bool = True) -> torch.Tensor:
        if frequencies.ndim != 1:
            raise ValueError("frequencies must have shape (frequency,)")
        coeff, tau = self.cir(normalize_delays=False)
        src = self.source
        valid = tau >= 0.0
        out = paths_cfr(coeff.reshape(src.num_rx * src.num_tx, -1, src.num_time_steps),
                        tau.reshape(src.num_rx * src.num_tx, -1), valid.reshape(src.num_rx * src.num_tx, -1),
                        frequencies, normalize_delays=normalize_delays)
        return out.reshape(src.num_rx, src.num_tx, src.num_time_steps, int(frequencies.shape[0]))


def endpoint_angles(direction: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(theta, phi) of a direction: theta = acos(z / |d|), phi = atan2(y, x)
    (reference path.py:879-884)."""
    norm = torch.linalg.vector_norm(direction, dim=-1)
    safe = torch.clamp(norm, min=torch.finfo(direction.dtype).tiny)
    theta = torch.acos(torch.clamp(direction[..., 2] / safe, -1.0, 1.0))
    phi = torch.atan2(direction[..., 1], direction[..., 0])
    return theta.to(torch.float32), phi.to(torch.float32)


def interaction_types(component_id: torch.Tensor, depth: torch.Tensor, width: int) -> torch.Tensor:
    """Per-slot event flags: reflection 1 on every bounce, diffraction 2 on
    slot 0, transmission 4 on every wall (reference propagation/topology.py:89-121)."""
    count = int(component_id.numel())
    out = torch.zeros((count, width), device=component_id.device, dtype=torch.int32)
    if count == 0 or width == 0:
        return out
    active = torch.arange(width, device=component_id.device).view(1, -1) < depth.long().view(-1, 1)
    out[active & (component_id.view(-1, 1) == 1)] = int(InteractionType.REFLECTION)
    out[(component_id == 2) & (depth > 0), 0] = int(InteractionType.DIFFRACTION)
    out[active & (component_id.view(-1, 1) == 5)] = int(InteractionType.TRANSMISSION)
    return out


def from_rows(rows: dict[str, Any], *, width: int, num_rx: int, num_tx: int, tx_positions: torch.Tensor,
              rx_positions: torch.Tensor, metadata: dict[str, Any] | None = None) -> PathResult:
    """Pack canonical export rows into a ``PathResult`` (reference
    ``from_evaluated_paths``, path.py:887-975): angles of departure from the
    first interaction (or the direct ray), of arrival from the last one; the
    coefficient is the unit-excitation receiver projection."""
    tx_id = rows["tx_id"].to(torch.int32)
    rx_id = rows["rx_id"].to(torch.int32)
    count = int(tx_id.shape[0])
    device = tx_id.device
    tx_for = tx_positions[tx_id.long()]
    rx_for = rx_positions[rx_id.long()]
    direct = rx_for - tx_for
    positions = rows["interaction_positions"].reshape(count, width, 3)
    depth = rows["depth"]
    if count and width:
        first = positions[:, 0]
        last = positions[torch.arange(count, device=device), torch.clamp(depth.long() - 1, min=0)]
        has = (depth > 0).view(-1, 1)
        departure = torch.where(has, first - tx_for, direct)
        arrival = torch.where(has, rx_for - last, direct)
    else:
        departure = arrival = direct
    theta_t, phi_t = endpoint_angles(departure)
    theta_r, phi_r = endpoint_angles(-arrival)
    comp = rows["component_id"]
    itype = interaction_types(comp, depth, width)
    objects = rows["primitive_sequence"].reshape(count, width).clone()
    diff = (comp == 2) & (depth > 0)
    if width:
        objects[diff, 0] = rows["edge_id"][diff].to(torch.int32)
    normals = rows["interaction_normals"].reshape(count, width, 3)
    normals = torch.where((itype == int(InteractionType.DIFFRACTION)).unsqueeze(-1) & ~torch.isfinite(normals),
                          torch.zeros_like(normals), normals)
    ragged = RaggedPathSoA.from_flat(
        num_rx=num_rx, num_rx_ant=1, num_tx=num_tx, num_tx_ant=1, rx_id=rx_id, tx_id=tx_id,
        field=rows["coefficient"].to(torch.complex64).unsqueeze(-1), delay_s=rows["delay_s"].to(torch.float32),
        theta_t=theta_t, phi_t=phi_t, theta_r=theta_r, phi_r=phi_r, interaction_type=itype,
        primitive_id=objects, material_id=rows["material_sequence"].reshape(count, width),
        position=positions.reshape(count, width, 3), normal=normals)
    meta = dict(metadata or {})
    meta.update({"coefficient_semantics": "unit_excitation_dimensionless_receiver_projection",
                 "coupled_coefficient_semantics": "unified_complex3_jones",
                 "interaction_geometry": "canonical_topology"})
    result = PathResult.from_ragged(ragged, metadata=meta)
    field_xyz = torch.zeros((*result.path_shape, 3), device=device, dtype=torch.complex64)
    field_direction = torch.zeros((*result.path_shape, 3), device=device, dtype=torch.float32)
    # rows are already in pair order, so the valid slots take them in row order
    field_xyz[result.valid] = rows["field_xyz"].to(torch.complex64)
    field_direction[result.valid] = rows["arrival_direction"].to(torch.float32)
    return replace(result, field_xyz=field_xyz, field_direction=field_direction)


def _metadata(config: Config, path_count: int, transmission_count: int) -> dict[str, Any]:
    depths = component_max_depth(config.components, chain_depth=config.max_depth, single_bounce_depth=1)
    status = {name: ("enabled" if name in config.components else "not_requested")
              for name in ("los", "reflection", "diffraction", "transmission", "scattering")}
    if "transmission" in config.components and transmission_count == 0:
        status["transmission"] = "enabled_no_paths"
    meta: dict[str, Any] = {
        "solver": "path",
        "device": "cuda",
        "path_count": path_count,
        "effective_max_depth": max(depths["los"], depths["reflection"], depths["diffraction"]),
        "component_max_depth": depths,
        "max_paths_per_pair": config.max_paths,
        "native_capabilities": {"path_native": True, "rayd_native": True, "reflection": True, "diffraction": True},
        "components": status,
        "kernel": {"primitive": "path_solver", "forward_launch_count": 1 if path_count else 0,
                   "scheduling_strategy": "fused_hip", "ad_status": config.ad_mode},
        "field_abi": "complex3_v1",
        "phase_convention": dict(PHASE_CONVENTION),
        "coefficient_semantics": "unit_excitation_dimensionless_receiver_projection",
        "coupled_paths": {"requested": config.coupled_paths, "geometry": "not_requested",
                          "coefficient": "not_requested"},
        "array_semantics": "single_element",
        "num_rx_ant": 1,
        "num_tx_ant": 1,
    }
    if "transmission" in config.components:
        meta["transmission"] = {"thin_sheet_straight_path_approximation": True, "group_delay": "geometric"}
    return meta


def _validate(config: Config) -> None:
    unsupported = set(config.components) - {"los", "reflection", "diffraction", "transmission"}
    if unsupported:
        raise NotImplementedError(f"components {sorted(unsupported)} are outside this build's hot path")
    if config.coupled_paths:
        raise NotImplementedError("coupled reflection-diffraction paths are outside this build's hot path")
    if config.isb_boundary_taper:
        raise NotImplementedError("isb_boundary_taper is outside this build's hot path")
    if config.ad_mode != "none":
        raise NotImplementedError("path.solve differentiates through deterministic.solve(ad_mode='vjp') in this "
                                  "build; explicit path rows are forward only")
    if config.max_depth < 1 and ("reflection" in config.components or "diffraction" in config.components):
        raise RuntimeError("requested scattering paths require max_depth >= 1")


def solve_compiled(compiled: CompiledScene, config: Config) -> PathResult:
    """Path solve of a compiled scene (all transmitters x all receivers)."""
    device = _native.require_gpu()
    _validate(config)
    ep = compiled.endpoints
    num_tx, num_rx = int(ep.tx_position.shape[0]), int(ep.rx_position.shape[0])
    if num_tx == 0 or num_rx == 0:
        raise RuntimeError("path solve requires at least one transmitter and one receiver")
    dscene = device_scene(compiled, device)
    bits = component_bits(config.components)
    depth = config.max_depth if bits & (2 | 4 | 8) else 0
    inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                 frequency_hz=compiled.frequency_hz)
    rows = engine.export_rows(dscene, inputs, width=max(depth, 1), max_paths_per_pair=config.max_paths)
    tx_pos = torch.as_tensor(ep.tx_position, dtype=torch.float32).to(device)
    rx_pos = torch.as_tensor(ep.rx_position, dtype=torch.float32).to(device)
    count = int(rows["tx_id"].shape[0])
    transmission = int((rows["component_id"] == 5).sum().item()) if count else 0
    meta = _metadata(config, count, transmission)
    result = from_rows(rows, width=max(depth, 1), num_rx=num_rx, num_tx=num_tx, tx_positions=tx_pos,
                       rx_positions=rx_pos, metadata=meta)
    if config.max_paths is not None:
        result = _truncate(result, config.max_paths)
    return result


def _truncate(result: PathResult, max_paths: int) -> PathResult:
    """Keep the first ``max_paths`` paths of every pair (canonical order)."""
    if result.max_num_paths <= max_paths:
        return result
    cut = lambda t: t[:, :, :, :, :max_paths].contiguous()  # noqa: E731
    return replace(result, a=cut(result.a), tau=cut(result.tau), theta_t=cut(result.theta_t),
                   phi_t=cut(result.phi_t), theta_r=cut(result.theta_r), phi_r=cut(result.phi_r),
                   valid=cut(result.valid), interaction_type=cut(result.interaction_type),
                   primitive_id=cut(result.primitive_id), material_id=cut(result.material_id),
                   position=cut(result.position), normal=cut(result.normal),
                   num_paths=torch.clamp(result.num_paths, max=max_paths), field_xyz=cut(result.field_xyz),
                   field_direction=cut(result.field_direction))


@dataclass
class BatchedPaths:
    """Explicit paths of transmitters with their own receiver sets, flat in
    canonical order: pairs are (local receiver, transmitter) with pair index
    rx * T + tx, rows of a pair contiguous (``pair_offset`` CSR over the
    ``pair_count`` pairs in that order) and ordered inside a pair as in
    ``solve``. ``a`` is the complex coefficient (unit transmit power), ``tau``
    the delay, theta / phi the departure (t) and arrival (r) angles."""

    rows: dict[str, Any]
    pair_offset: torch.Tensor   # [pairs + 1] int64
    pair_tx: torch.Tensor       # [pairs] int64
    pair_rx: torch.Tensor       # [pairs] int64 (index in the transmitter's receiver set)
    a: torch.Tensor
    tau: torch.Tensor
    theta_t: torch.Tensor
    phi_t: torch.Tensor
    theta_r: torch.Tensor
    phi_r: torch.Tensor

    @property
    def path_count(self) -> int:
        return int(self.tau.shape[0])

    @property
    def pair_count(self) -> int:
        return int(self.pair_tx.shape[0])

    def padded(self) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """(a [pairs, P, 1], tau [pairs, P], valid [pairs, P]) padded to the
        largest path count of a pair."""
        device = self.tau.device
        counts = (self.pair_offset[1:] - self.pair_offset[:-1])
        width = int(counts.max().item()) if counts.numel() else 0
        pairs = self.pair_count
        a = torch.zeros((pairs, max(width, 1), 1), dtype=torch.complex64, device=device)
        tau = torch.zeros((pairs, max(width, 1)), dtype=torch.float32, device=device)
        valid = torch.zeros((pairs, max(width, 1)), dtype=torch.bool, device=device)
        if self.path_count:
            pair_of_row = torch.repeat_interleave(torch.arange(pairs, device=device), counts)
            slot = torch.arange(self.path_count, device=device) - self.pair_offset[:-1][pair_of_row]
            a[pair_of_row, slot, 0] = self.a
            tau[pair_of_row, slot] = self.tau
            valid[pair_of_row, slot] = True
        return a, tau, valid

    def cfr(self, frequencies: torch.Tensor, *, normalize_delays: bool = False) -> torch.Tensor:
        """[pairs, F] c64 channel frequency response of every pair
        (``wc_paths_cfr``: sum over its paths of a exp(-j 2 pi f (tau - tau0)),
        tau0 = 0 unless ``normalize_delays``)."""
        a, tau, valid = self.padded()
        return paths_cfr(a, tau, valid, frequencies, normalize_delays=normalize_delays)[:, 0, :]


def solve_batched(compiled: CompiledScene, config: Config, tx_positions: np.ndarray, rx_positions: np.ndarray,
                  rx_count: np.ndarray, *, dscene=None, inputs=None) -> BatchedPaths:
    """Explicit paths of many transmitters, each with its own receiver set
    (the consecutive sets of sizes ``rx_count`` in ``rx_positions``), in one
    batched launch set: the row count, the rows and their canonical order for
    every (transmitter, receiver) pair of the batch (the reference solves the
    full transmitter x receiver product, path.py:1551-1648)."""
    device = _native.require_gpu()
    _validate(config)
    if dscene is None:
        dscene = device_scene(compiled, device)
    bits = component_bits(config.components)
    depth = config.max_depth if bits & (2 | 4 | 8) else 0
    tx_positions = np.ascontiguousarray(tx_positions, dtype=np.float32).reshape(-1, 3)
    rx_positions = np.ascontiguousarray(rx_positions, dtype=np.float32).reshape(-1, 3)
    t = tx_positions.shape[0]
    pol = lambda n: np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (n, 1))  # noqa: E731
    if inputs is None:
        inputs = engine.solve_inputs(compiled, dscene, components=bits, max_depth=depth, coherent=True,
                                     frequency_hz=compiled.frequency_hz, tx_pos=tx_positions, tx_pol=pol(t),
                                     tx_power=np.ones(t, np.float32), rx_pos=rx_positions,
                                     rx_pol=pol(rx_positions.shape[0]), rx_count=rx_count)
    rows = engine.export_rows(dscene, inputs, width=max(depth, 1), max_paths_per_pair=config.max_paths)
    offset = rows.pop("pair_offset")
    cells = rows.pop("pair_cell")
    rows.pop("dropped")
    # transmitter and receiver of each canonical pair, from the cell layout
    cell_offset = inputs.cell_offset.to(torch.int64)
    pair_tx = torch.searchsorted(cell_offset, cells, right=True) - 1
    pair_rx = cells - cell_offset[pair_tx]
    theta_t, phi_t = endpoint_angles(rows["departure_direction"])
    theta_r, phi_r = endpoint_angles(-rows["arrival_direction"])
    return BatchedPaths(rows=rows, pair_offset=offset, pair_tx=pair_tx, pair_rx=pair_rx,
                        a=rows["coefficient"].to(torch.complex64), tau=rows["delay_s"].to(torch.float32),
                        theta_t=theta_t, phi_t=phi_t, theta_r=theta_r, phi_r=phi_r)


def solve(scene, config: Config, *, reference_frequency_hz) -> PathResult:  # type: ignore[no-untyped-def]
    """Solve canonical paths and pack them per (receiver, transmitter) pair."""

    compiled = compile_scene(scene, reference_frequency_hz=reference_frequency_hz)
    return solve_compiled(compiled, config)


__all__ = ["BatchedPaths", "Config", "InteractionType", "PathResult", "RaggedPathSoA", "BeamformedPathResult",
           "endpoint_angles", "from_rows", "paths_cfr", "solve", "solve_batched"]
