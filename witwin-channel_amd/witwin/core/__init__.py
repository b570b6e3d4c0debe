"""Minimal scene model with the fields the propagation hot path reads.

The reference solvers consume scenes built from the external ``witwin==0.4.0``
package (``witwin.core``: Scene, Mesh, Structure, PhysicalMaterial,
MaterialLayer, AntennaState, ReceiverGrid). That package is not part of this
build, so this module provides the same constructor signatures and attribute
names for the subset the deterministic/path pipeline reads
(reference call sites: tests/support/core_world.py:12-90,
witwin/channel/scene/compiler.py:967-1036, witwin/channel/scene/endpoints.py:289-326).
Everything here is plain host-side data; no geometry processing happens here.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Any, Sequence

import torch

__all__ = [
    "AntennaState",
    "MaterialLayer",
    "Mesh",
    "PhysicalMaterial",
    "ReceiverGrid",
    "Scene",
    "SceneSnapshot",
    "Structure",
    "SurfaceRoughness",
]

_ID_COUNTER = itertools.count(1)


def new_antenna_id() -> int:
    return next(_ID_COUNTER)


def _vec3(value: Any, name: str) -> torch.Tensor:
    tensor = torch.as_tensor(value, dtype=torch.float32)
    if tensor.shape != (3,):
        raise ValueError(f"{name} must be a 3-vector")
    return tensor


@dataclass(frozen=True)
class MaterialLayer:
    """One homogeneous layer of a thin-sheet wall (entry side first)."""

    thickness_m: float = 0.1
    eps_r: float | complex = 1.0
    sigma_e: float = 0.0
    mu_r: float = 1.0
    dispersion: Any = None

    def __post_init__(self) -> None:
        if float(self.thickness_m) < 0.0:
            raise ValueError("thickness_m must be non-negative")


@dataclass(frozen=True)
class SurfaceRoughness:
    rms_height_m: float
    correlation_length_x_m: float
    correlation_length_y_m: float
    principal_axis_rad: float = 0.0


@dataclass(frozen=True)
class PhysicalMaterial:
    """Surface material. ``layers`` (if given) overrides the scalar first layer.

    ``conductor_model="perfect"`` marks a PEC, which the solvers realize through
    an effective conductivity of 1e9 S/m (reference materials.py:29-31).
    """

    eps_r: float | complex = 1.0
    sigma_e: float = 0.0
    mu_r: float = 1.0
    thickness_m: float = 0.1
    gain: float = 1.0
    layers: tuple[MaterialLayer, ...] | None = None
    name: str | None = None
    conductor_model: str | None = None
    geometry_mode: str = "thin_sheet"
    roughness_front: SurfaceRoughness | None = None
    roughness_back: SurfaceRoughness | None = None
    scattering_coefficient: float = 0.0
    xpd_coefficient: float = 0.0
    dispersion: Any = None

    def __post_init__(self) -> None:
        if self.layers is not None:
            object.__setattr__(self, "layers", tuple(self.layers))
            if not self.layers:
                raise ValueError("layers must be non-empty when given")
        if self.geometry_mode not in ("thin_sheet", "volumetric"):
            raise ValueError("geometry_mode must be 'thin_sheet' or 'volumetric'")

    @classmethod
    def perfect_conductor(cls, *, name: str | None = "pec", thickness_m: float = 0.1) -> "PhysicalMaterial":
        return cls(eps_r=1.0, sigma_e=0.0, mu_r=1.0, thickness_m=thickness_m,
                   name=name, conductor_model="perfect")

    def layer_rows(self) -> tuple[MaterialLayer, ...]:
        if self.layers:
            return self.layers
        return (MaterialLayer(thickness_m=self.thickness_m, eps_r=self.eps_r,
                              sigma_e=self.sigma_e, mu_r=self.mu_r),)


class Mesh:
    """Triangle mesh: ``vertices`` (V, 3) float32 and ``faces`` (F, 3) int."""

    def __init__(self, vertices: Any, faces: Any, *, recenter: bool = False,
                 fill_mode: str = "surface", topology_diagnostics: bool = False) -> None:
        v = torch.as_tensor(vertices, dtype=torch.float32)
        f = torch.as_tensor(faces).to(torch.int64)
        if v.ndim != 2 or v.shape[1] != 3:
            raise ValueError("vertices must have shape (V, 3)")
        if f.ndim != 2 or f.shape[1] != 3:
            raise ValueError("faces must have shape (F, 3)")
        if f.numel() and (int(f.min()) < 0 or int(f.max()) >= v.shape[0]):
            raise ValueError("faces index outside vertices")
        if recenter:
            v = v - v.mean(dim=0, keepdim=True)
        self.vertices = v
        self.faces = f
        self.fill_mode = fill_mode


class Structure:
    """A mesh with one material and a surface id (used by coplanar grouping)."""

    def __init__(self, geometry: Mesh, material: PhysicalMaterial, *, name: str | None = None,
                 surface_id: int | None = None, phase_screen: Any = None,
                 uv: Any = None, face_uv: Any = None) -> None:
        self.geometry = geometry
        self.material = material
        self.name = name
        self.surface_id = surface_id
        self.phase_screen = phase_screen
        self.uv = uv
        self.face_uv = face_uv


class AntennaState:
    """A transmitter (role ``"tx"``) or receiver point (role ``"rx"``).

    ``polarization`` defaults to world +z; ``power_w`` defaults to 1 W.
    """

    def __init__(self, antenna_id: int, role: str, position: Any, *, orientation: Any = None,
                 polarization: Any = None, element_positions: Any = None, weights: Any = None,
                 synthetic_array: bool = True, pattern: str = "isotropic",
                 power_w: float | None = 1.0) -> None:
        if role not in ("tx", "rx"):
            raise ValueError("role must be 'tx' or 'rx'")
        if pattern != "isotropic":
            raise NotImplementedError("only the isotropic pattern is supported")
        self.antenna_id = antenna_id
        self.role = role
        self.position = torch.as_tensor(position, dtype=torch.float32)
        self.orientation = orientation
        self.polarization = None if polarization is None else _vec3(polarization, "polarization")
        self.element_positions = element_positions
        self.weights = weights
        self.synthetic_array = synthetic_array
        self.pattern = pattern
        self.power_w = power_w


class ReceiverGrid(AntennaState):
    """Planar receiver grid: point(row, col) = origin + x_axis*row*s0 + y_axis*col*s1.

    Flat receiver index = row * cols + col (reference los.cu:391-420).
    """

    def __init__(self, antenna_id: int, origin: Any, x_axis: Any, y_axis: Any,
                 shape: Sequence[int], spacing: Any, **kwargs: Any) -> None:
        super().__init__(antenna_id, "rx", origin, **kwargs)
        self.origin = _vec3(origin, "origin")
        self.x_axis = _vec3(x_axis, "x_axis")
        self.y_axis = _vec3(y_axis, "y_axis")
        rows, cols = (int(shape[0]), int(shape[1]))
        if rows <= 0 or cols <= 0:
            raise ValueError("grid shape must be positive")
        self.shape = (rows, cols)
        if isinstance(spacing, (int, float)):
            spacing = (float(spacing), float(spacing))
        self.spacing = (float(spacing[0]), float(spacing[1]))

    def points(self) -> torch.Tensor:
        rows, cols = self.shape
        r = torch.arange(rows, dtype=torch.float32) * self.spacing[0]
        c = torch.arange(cols, dtype=torch.float32) * self.spacing[1]
        pts = (self.origin[None, None, :]
               + self.x_axis[None, None, :] * r[:, None, None]
               + self.y_axis[None, None, :] * c[None, :, None])
        return pts.reshape(rows, cols, 3)


@dataclass
class Scene:
    structures: list[Structure] = field(default_factory=list)
    endpoints: list[AntennaState] = field(default_factory=list)
    versions: dict[str, int] = field(default_factory=dict)
    metadata: dict[str, Any] = field(default_factory=dict)

    @property
    def transmitters(self) -> tuple[AntennaState, ...]:
        return tuple(e for e in self.endpoints if e.role == "tx")

    @property
    def receivers(self) -> tuple[AntennaState, ...]:
        return tuple(e for e in self.endpoints if e.role == "rx")


SceneSnapshot = Scene
