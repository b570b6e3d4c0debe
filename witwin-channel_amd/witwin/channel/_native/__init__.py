"""ctypes binding of libwcmi.so, the gfx950 HIP library behind the hot path.

The library exports a plain C ABI (declared in include/wcmi.h at the repository
root). This module mirrors its structs and wraps each entry point so a non-zero
status raises ``RuntimeError`` with ``wc_last_error()``. There is no fallback:
if the library is missing or fails to load, every solver raises.
(Reference loader contract: witwin/channel/runtime.py:552-579 — missing native
symbols raise instead of falling back.)
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Any

import torch

_LIB_PATH = Path(__file__).resolve().parent / "libwcmi.so"
_lib: ctypes.CDLL | None = None

c_int = ctypes.c_int
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
c_double = ctypes.c_double
c_double_p = ctypes.POINTER(ctypes.c_double)


class NativeLibraryError(RuntimeError):
    """libwcmi.so is missing or does not export a required symbol."""


class WcEdge(ctypes.Structure):
    _fields_ = [(name, ctypes.c_float) for name in (
        "ax", "ay", "az", "dx", "dy", "dz", "length", "n", "t0x", "t0y", "t0z", "n0x", "n0y", "n0z")] + [
        (name, ctypes.c_int) for name in ("face0", "face1", "group0", "group1", "material0", "material1")]


class WcScene(ctypes.Structure):
    _fields_ = [
        ("nodes", c_void_p), ("tris", c_void_p), ("tri_count", c_int), ("node_count", c_int),
        ("faces", c_void_p), ("face_group", c_void_p), ("face_material", c_void_p), ("face_count", c_int),
        ("group_plane", c_void_p), ("group_face_offset", c_void_p), ("group_faces", c_void_p),
        ("group_material", c_void_p), ("group_count", c_int),
        ("mat_eps", c_void_p), ("mat_sigma", c_void_p), ("mat_mu", c_void_p), ("mat_gain", c_void_p),
        ("mat_thickness", c_void_p), ("mat_layer_offset", c_void_p), ("mat_layer_count", c_void_p),
        ("layer_thickness", c_void_p), ("layer_eps", c_void_p), ("layer_sigma", c_void_p),
        ("layer_mu", c_void_p), ("material_count", c_int),
        ("edges", c_void_p), ("edge_count", c_int),
    ]


class WcSolve(ctypes.Structure):
    _fields_ = [
        ("tx_pos", c_void_p), ("tx_pol", c_void_p), ("tx_power", c_void_p), ("tx_count", c_int),
        ("rx_pos", c_void_p), ("rx_pol", c_void_p), ("rx_pol_stride", c_int), ("cell_offset", c_void_p),
        ("rx_shared", c_int), ("total_cells", c_int),
        ("chain_offset", c_void_p), ("chain_groups", c_void_p), ("chain_images", c_void_p),
        ("frequency_hz", c_double), ("components", c_int), ("max_depth", c_int), ("coherent", c_int),
        ("lanes_per_cell", c_int), ("cell_order", c_void_p), ("max_tx_chains", c_int),
        ("super_first_tile", c_void_p), ("tile_super", c_void_p), ("super_count", c_int),
        ("visibility_mode", c_int), ("chain_cones", c_void_p),
    ]


class WcRadiomapOut(ctypes.Structure):
    _fields_ = [
        ("power_total", c_void_p), ("field_total", c_void_p), ("component_power", c_void_p),
        ("component_field", c_void_p), ("diffraction_vector", c_void_p), ("path_count", c_void_p),
        ("topology", c_void_p), ("topology_capacity", c_int), ("failure", c_void_p), ("workspace", c_void_p),
    ]


class WcPathRows(ctypes.Structure):
    _fields_ = [("width", c_int)] + [(name, c_void_p) for name in (
        "tx_id", "rx_id", "depth", "component_id", "primitive_id", "edge_id", "material_id", "path_length_m",
        "delay_s", "path_gain", "primitive_sequence", "material_sequence", "interaction_positions",
        "interaction_normals", "path_field", "coefficient", "field_xyz", "arrival_direction",
        "departure_direction")]


class WcAccumRows(ctypes.Structure):
    _fields_ = [("count", c_int), ("num_tx", c_int), ("num_rx", c_int), ("coherent", c_int),
                ("scattering_coherent", c_int), ("dtype", c_int), ("valid", c_void_p), ("tx_id", c_void_p),
                ("rx_id", c_void_p), ("component_id", c_void_p), ("path_gain", c_void_p), ("field_real", c_void_p),
                ("field_imag", c_void_p)]


class WcAccumOut(ctypes.Structure):
    _fields_ = [("power_total", c_void_p), ("field_total_real", c_void_p), ("field_total_imag", c_void_p),
                ("component_power", c_void_p), ("component_field_real", c_void_p),
                ("component_field_imag", c_void_p)]


class WcTangentIn(ctypes.Structure):
    _fields_ = [("d_eps_r", c_void_p), ("d_sigma_e", c_void_p), ("d_frequency", c_float), ("d_tx_pos", c_void_p)]


class WcTangentOut(ctypes.Structure):
    _fields_ = [("d_power_total", c_void_p), ("d_field_total", c_void_p), ("d_component_power", c_void_p),
                ("d_component_field", c_void_p), ("d_diffraction_power", c_void_p)]


class WcPlanLevels(ctypes.Structure):
    _fields_ = [("prefix_parent", c_void_p * 5), ("prefix_group", c_void_p * 5), ("prefix_tx", c_void_p * 5)]


class WcLayerTables(ctypes.Structure):
    _fields_ = [("layer_offset", c_void_p), ("layer_count", c_void_p), ("layer_thickness", c_void_p),
                ("layer_eps", c_void_p), ("layer_sigma", c_void_p), ("layer_mu", c_void_p),
                ("material_count", c_int), ("layer_total", c_int)]


_FIELD_ROW_NAMES = ("field_vector", "coefficient", "path_field", "path_gain", "path_length_m", "delay_s",
                    "direction")


class WcFieldRowsOut(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in _FIELD_ROW_NAMES]


class WcFieldRowsGrad(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in _FIELD_ROW_NAMES]


class WcReflectionRows(ctypes.Structure):
    _fields_ = [("count", c_int), ("depth", c_int), ("frequency_hz", ctypes.c_double)] + [
        (name, c_void_p) for name in ("source", "target", "interaction_positions", "interaction_normals",
                                      "tx_power", "tx_polarization", "rx_polarization", "eps_r", "sigma_e",
                                      "mu_r", "gain", "thickness")]


class WcReflectionRowsGradOut(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in (
        "grad_source", "grad_target", "grad_interaction_positions", "grad_interaction_normals", "grad_eps_r",
        "grad_sigma_e", "grad_gain", "grad_thickness", "grad_frequency")]


class WcTransmissionRows(ctypes.Structure):
    _fields_ = [("count", c_int), ("depth", c_int), ("frequency_hz", ctypes.c_double)] + [
        (name, c_void_p) for name in ("path_valid", "interaction_valid", "interaction_material_id", "source",
                                      "target", "interaction_normals", "tx_power", "tx_polarization",
                                      "rx_polarization")] + [("layers", WcLayerTables)]


LAYER_STACK_FIELDS = ("r_te_real", "r_te_imag", "r_tm_real", "r_tm_imag", "t_te_real", "t_te_imag", "t_tm_real",
                      "t_tm_imag", "cap_R_te", "cap_R_tm", "cap_T_te", "cap_T_tm")


class WcLayerStackOut(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in LAYER_STACK_FIELDS]


class WcCoupledOut(ctypes.Structure):
    _fields_ = [("field", c_void_p), ("power", c_void_p), ("count", c_void_p), ("count_dd", c_void_p)]


class WcFixedRows(ctypes.Structure):
    _fields_ = [
        ("count", c_int), ("width", c_int), ("component", c_void_p), ("depth", c_void_p), ("faces", c_void_p),
        ("source", c_void_p), ("target", c_void_p), ("tx_power", c_void_p), ("tx_polarization", c_void_p),
        ("rx_polarization", c_void_p), ("valid", c_void_p), ("path_field", c_void_p), ("field_xyz", c_void_p),
        ("path_length", c_void_p), ("delay", c_void_p), ("direction", c_void_p), ("positions", c_void_p),
        ("normals", c_void_p),
    ]


class WcRefit(ctypes.Structure):
    _fields_ = [
        ("vertices", c_void_p), ("faces", c_void_p), ("group_face", c_void_p), ("edge_vertices", c_void_p),
        ("node_order", c_void_p), ("level_offset_host", c_void_p), ("levels", c_int), ("failure", c_void_p),
    ]


class WcAdjoint(ctypes.Structure):
    _fields_ = [
        ("grad_power_total", c_void_p), ("grad_field_total", c_void_p), ("grad_component_power", c_void_p),
        ("grad_component_field", c_void_p), ("grad_diffraction_power", c_void_p), ("power_total", c_void_p),
        ("field_total", c_void_p), ("component_field", c_void_p), ("diffraction_vector", c_void_p),
        ("topology", c_void_p), ("path_count", c_void_p), ("topology_capacity", c_int), ("partials", c_void_p),
    ]


_SIGNATURES: dict[str, tuple[Any, list[Any]]] = {
    "wc_last_error": (c_char_p, []),
    "wc_clear_error": (None, []),
    "wc_abi_version": (c_int, []),
    "wc_build_target": (c_char_p, []),
    "wc_bvh_workspace_bytes": (c_size_t, [c_int]),
    "wc_bvh_node_count": (c_int, [c_int]),
    "wc_bvh_widen": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "wc_bvh_build_sah": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "wc_bvh_build": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_size_t, c_void_p]),
    "wc_radiomap_workspace_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    "wc_radiomap_forward": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve),
                                    ctypes.POINTER(WcRadiomapOut), c_void_p, c_int, c_void_p, c_void_p]),
    "wc_radiomap_pass": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve),
                                 ctypes.POINTER(WcRadiomapOut), c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "wc_radiomap_vjp": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve), ctypes.POINTER(WcAdjoint),
                                c_void_p, c_int, c_void_p]),
    "wc_radiomap_vjp_phase": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve), ctypes.POINTER(WcAdjoint),
                                      c_void_p, c_int, c_int, c_void_p]),
    "wc_radiomap_jvp": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve), ctypes.POINTER(WcAdjoint),
                                ctypes.POINTER(WcTangentIn), ctypes.POINTER(WcTangentOut), c_void_p]),
    "wc_vjp_partial_stride": (c_int, [c_int]),
    "wc_vjp_partial_rows": (c_int, [c_int]),
    "wc_vjp_reduce": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "wc_scene_refit": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcRefit), c_void_p]),
    "wc_fixed_rows_evaluate": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcFixedRows), ctypes.c_double,
                                       c_void_p]),
    "wc_radiomap_vjp_geometry": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve),
                                         ctypes.POINTER(WcAdjoint), c_void_p, c_int, c_void_p, c_void_p,
                                         c_void_p]),
    "wc_geometry_vertex_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                        c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "wc_paths_count": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve), c_void_p, c_void_p]),
    "wc_paths_emit": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve), c_void_p, c_void_p,
                              ctypes.POINTER(WcPathRows), c_void_p]),
    "wc_paths_canonical_order": (c_int, [ctypes.POINTER(WcPathRows), c_void_p, c_int, c_void_p, c_void_p,
                                         c_void_p]),
    "wc_paths_compact_workspace_bytes": (c_size_t, [c_int]),
    "wc_plan_workspace_bytes": (c_size_t, [c_int, c_int]),
    "wc_chain_cones": (c_int, [ctypes.POINTER(WcScene), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                               c_void_p, c_void_p]),
    "wc_plan_level": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "wc_plan_assign": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int] + [c_void_p] * 7),
    "wc_plan_table": (c_int, [ctypes.POINTER(WcPlanLevels), c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p]),
    "wc_plan_exhaustive": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "wc_paths_pair_order": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "wc_paths_row_offsets": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "wc_paths_compact": (c_int, [ctypes.POINTER(WcPathRows), c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "wc_paths_cfr": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, ctypes.c_double,
                             ctypes.c_double, c_int, c_int, c_void_p, c_void_p]),
    "wc_utd_special_eval": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "wc_field_free_space": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_double,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p]),
    "wc_path_los_export": (c_int, [c_int, c_int] + [c_void_p] * 4 + [ctypes.c_double] + [c_void_p] * 7),
    "wc_field_project_complex3": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "wc_field_reflection_sequence": (c_int, [c_int, c_int] + [c_void_p] * 12 + [ctypes.c_double] + [c_void_p] * 8),
    "wc_coupled_candidates_per_cell": (ctypes.c_longlong, [c_int, c_int, c_int]),
    "wc_coupled_workspace_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "wc_coupled_paths": (c_int, [ctypes.POINTER(WcScene), ctypes.POINTER(WcSolve), c_int, c_int, c_int,
                                 ctypes.POINTER(WcCoupledOut),
                                 c_void_p, c_size_t, c_void_p]),
    "wc_field_rows_workspace_bytes": (c_size_t, [c_int, c_int]),
    "wc_field_free_space_jvp": (c_int, [c_int] + [c_void_p] * 5 + [ctypes.c_double, c_float, c_void_p, c_void_p,
                                        ctypes.POINTER(WcFieldRowsOut), c_void_p]),
    "wc_field_free_space_vjp": (c_int, [c_int] + [c_void_p] * 5 + [ctypes.c_double, ctypes.POINTER(WcFieldRowsGrad)]
                                + [c_void_p] * 5),
    "wc_field_reflection_sequence_jvp": (c_int, [ctypes.POINTER(WcReflectionRows), ctypes.POINTER(WcReflectionRows),
                                                 c_float, ctypes.POINTER(WcFieldRowsOut), c_void_p]),
    "wc_field_reflection_sequence_vjp": (c_int, [ctypes.POINTER(WcReflectionRows), ctypes.POINTER(WcFieldRowsGrad),
                                                 ctypes.POINTER(WcReflectionRowsGradOut), c_void_p, c_void_p]),
    "wc_field_transmission_sequence": (c_int, [ctypes.POINTER(WcTransmissionRows), ctypes.POINTER(WcFieldRowsOut),
                                               c_void_p]),
    "wc_field_transmission_sequence_jvp": (c_int, [ctypes.POINTER(WcTransmissionRows)] + [c_void_p] * 6
                                           + [c_float, ctypes.POINTER(WcFieldRowsOut), c_void_p]),
    "wc_field_transmission_sequence_vjp": (c_int, [ctypes.POINTER(WcTransmissionRows),
                                                   ctypes.POINTER(WcFieldRowsGrad)] + [c_void_p] * 9),
    "wc_em_layer_stack_eval": (c_int, [c_int, c_void_p, c_void_p, ctypes.POINTER(WcLayerTables), ctypes.c_double,
                                       ctypes.POINTER(WcLayerStackOut), c_void_p]),
    "wc_em_layer_stack_jvp": (c_int, [c_int, c_void_p, c_void_p, ctypes.POINTER(WcLayerTables), ctypes.c_double]
                              + [c_void_p] * 4 + [c_float, ctypes.POINTER(WcLayerStackOut), c_void_p]),
    "wc_em_layer_stack_vjp": (c_int, [c_int, c_void_p, c_void_p, ctypes.POINTER(WcLayerTables), ctypes.c_double,
                                      ctypes.POINTER(WcLayerStackOut)] + [c_void_p] * 7),
    "wc_accumulate_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "wc_accumulate_flat": (c_int, [c_int] + [c_void_p] * 7 + [c_int, c_int, c_int] + [c_void_p] * 5
                           + [c_size_t, c_void_p]),
    "wc_accumulate_noalloc_max_bytes": (c_size_t, []),
    "wc_accumulate_rows_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "wc_accumulate_rows": (c_int, [ctypes.POINTER(WcAccumRows), ctypes.POINTER(WcAccumOut), c_void_p, c_void_p]),
    "wc_accumulate_rows_jvp": (c_int, [ctypes.POINTER(WcAccumRows), ctypes.POINTER(WcAccumOut),
                                       ctypes.POINTER(WcAccumOut), c_void_p, c_void_p]),
    "wc_accumulate_rows_backward": (c_int, [ctypes.POINTER(WcAccumRows), ctypes.POINTER(WcAccumOut),
                                            ctypes.POINTER(WcAccumOut), c_void_p, c_void_p, c_void_p, c_void_p]),
    "wc_accumulate_flat_noalloc": (c_int, [c_int] + [c_void_p] * 7 + [c_int, c_int, c_int] + [c_void_p] * 5),
    "wc_visibility": (c_int, [ctypes.POINTER(WcScene), c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "wc_discover_chains": (c_int, [ctypes.POINTER(WcScene), c_void_p, c_int, c_int, c_int, c_void_p,
                                   c_void_p]),
    "wc_chain_images": (c_int, [ctypes.POINTER(WcScene), c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_void_p, c_void_p]),
}


# int-returning entry points whose result is a value, not a status
_VALUE_RETURNING = frozenset({"wc_abi_version", "wc_bvh_node_count", "wc_vjp_partial_stride",
                              "wc_vjp_partial_rows"})


def library_path() -> Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load libwcmi.so once; raise NativeLibraryError when it is absent."""

    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("WCMI_LIBRARY", str(_LIB_PATH)))
    if not path.exists():
        raise NativeLibraryError(
            f"libwcmi.so not found at {path}; build it with "
            "`python witwin-channel_amd/build_native.py` (there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(path))
    for name, (restype, argtypes) in _SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            raise NativeLibraryError(f"libwcmi.so does not export {name}")
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGNATURES)


def call(name: str, *args: Any) -> Any:
    lib = load()
    fn = getattr(lib, name)
    status = fn(*args)
    if fn.restype is c_int and name not in _VALUE_RETURNING and status != 0:
        message = lib.wc_last_error().decode("utf-8", "replace")
        raise RuntimeError(f"{name} failed: {message}")
    return status


def ptr(tensor: torch.Tensor | None) -> int | None:
    if tensor is None:
        return None
    if not tensor.is_contiguous():
        raise ValueError("native calls require contiguous tensors")
    return tensor.data_ptr() if tensor.numel() else None


def stream_handle(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("witwin.channel requires a ROCm GPU (torch.cuda.is_available() is False)")
    load()
    return torch.device("cuda", torch.cuda.current_device())
