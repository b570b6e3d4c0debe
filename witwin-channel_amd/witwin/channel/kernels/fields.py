"""Row-level field operators (reference witwin/channel/kernels/fields.py):
free space (:110), receiver projection (:149), reflection sequences (:180)
and straight transmission through layered walls (:245), each with its
reverse-mode companion (``*_backward``, :551/:607/:710), forward-mode
companion (``*_jvp``, :583/:665/:771) and differentiable ``*_ad`` form
(:1255/:1522/:1786). Same argument meaning, output keys and shapes as the
reference; every operator runs in libwcmi.so (csrc/wc_fields.hip forward
free-space / reflection / projection, csrc/wc_rows.hip transmission and all
tangent / adjoint kernels). There is no CPU path.

Output schema of the enumerated-field operators: ``field_vector`` [N, 3]
complex64, ``coefficient`` / ``path_field`` [N] complex64, ``path_gain``,
``path_length_m``, ``delay_s`` [N] float32, ``direction`` [N, 3] float32.
Complex cotangents follow torch's convention (the gradient of a real loss
with respect to a + ib is dL/da + i dL/db)."""

from __future__ import annotations

import ctypes

import torch

from witwin.channel import _native

FIELD_OUTPUT_FIELDS = ("field_vector", "coefficient", "path_field", "path_gain", "path_length_m", "delay_s",
                       "direction")
FIELD_TANGENT_FIELDS = FIELD_OUTPUT_FIELDS[:6]
MAX_TRANSMISSION_DEPTH = 8
MAX_REFLECTION_DEPTH = 5
MAX_LAYERS = 8


# ----------------------------------------------------------------------------- validation
def _check(name, t, dtype, ndim, trailing=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}")
    if t.dim() != ndim:
        raise ValueError(f"{name} must have {ndim} dimensions")
    if trailing is not None and tuple(t.shape[-len(trailing):]) != tuple(trailing):
        raise ValueError(f"{name} must have trailing shape {tuple(trailing)}")


def _check_positive_frequency(frequency_hz):
    if not float(frequency_hz) > 0.0:
        raise ValueError("frequency_hz must be positive")


def _check_endpoints(count, source, target, tx_power, tx_polarization, rx_polarization, what):
    for name, value in (("source", source), ("target", target), ("tx_polarization", tx_polarization),
                        ("rx_polarization", rx_polarization)):
        _check(name, value, torch.float32, 2, (3,))
    _check("tx_power", tx_power, torch.float32, 1)
    if any(int(v.shape[0]) != count for v in (target, tx_power, tx_polarization, rx_polarization)):
        raise ValueError(f"{what} tensors must have matching rows")


def validate_layer_csr(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                       device) -> None:
    """CSR layer tables (reference kernels/materials.py:42-70)."""
    _check("layer_offset", layer_offset, torch.int32, 1)
    _check("layer_count", layer_count, torch.int32, 1)
    if layer_count.shape != layer_offset.shape:
        raise ValueError("layer_count must match layer_offset length")
    for name, t in (("layer_thickness_m", layer_thickness_m), ("layer_eps_r", layer_eps_r),
                    ("layer_sigma_e", layer_sigma_e), ("layer_mu_r", layer_mu_r)):
        _check(name, t, torch.float32, 1)
        if t.shape != layer_thickness_m.shape:
            raise ValueError(f"{name} must match layer_thickness_m length")
    for name, t in (("layer_offset", layer_offset), ("layer_count", layer_count),
                    ("layer_thickness_m", layer_thickness_m), ("layer_eps_r", layer_eps_r),
                    ("layer_sigma_e", layer_sigma_e), ("layer_mu_r", layer_mu_r)):
        if t.device != device:
            raise ValueError(f"{name} must share the op device")


def layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r, keep):
    """WcLayerTables over the CSR tensors; ``keep`` (a list) holds the
    contiguous views alive for the native call."""
    if layer_count.numel() and int(layer_count.max()) > MAX_LAYERS:
        raise ValueError(f"a material may have at most {MAX_LAYERS} layers")
    k = keep.append
    views = [t.detach().contiguous() for t in (layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                               layer_sigma_e, layer_mu_r)]
    for v in views:
        k(v)
    p = _native.ptr
    return _native.WcLayerTables(
        layer_offset=p(views[0]), layer_count=p(views[1]), layer_thickness=p(views[2]), layer_eps=p(views[3]),
        layer_sigma=p(views[4]), layer_mu=p(views[5]), material_count=int(layer_offset.shape[0]),
        layer_total=int(layer_thickness_m.shape[0]))


# ----------------------------------------------------------------------------- buffers
class _Keep(list):
    """Pointer of a contiguous view of t, kept alive until the native call ends."""

    def __call__(self, t):
        if t is None:
            return None
        t = t.contiguous()
        self.append(t)
        return _native.ptr(t)


def _alloc_rows(count, device, with_direction=True) -> dict[str, torch.Tensor]:
    out = {
        "field_vector": torch.empty((count, 3), dtype=torch.complex64, device=device),
        "coefficient": torch.empty(count, dtype=torch.complex64, device=device),
        "path_field": torch.empty(count, dtype=torch.complex64, device=device),
        "path_gain": torch.empty(count, dtype=torch.float32, device=device),
        "path_length_m": torch.empty(count, dtype=torch.float32, device=device),
        "delay_s": torch.empty(count, dtype=torch.float32, device=device),
    }
    if with_direction:
        out["direction"] = torch.empty((count, 3), dtype=torch.float32, device=device)
    return out


def _rows_out_view(out):
    return _native.WcFieldRowsOut(**{name: _native.ptr(out.get(name)) for name in FIELD_OUTPUT_FIELDS})


def _grad_view(grads: dict) -> tuple[_native.WcFieldRowsGrad, list]:
    """Cotangents as float views (complex -> interleaved re/im); keeps them alive."""
    keep = []
    fields = {}
    for name in FIELD_OUTPUT_FIELDS:
        g = grads.get(name)
        if g is None:
            fields[name] = None
            continue
        g = g.detach()
        if g.is_complex():
            g = torch.view_as_real(g.to(torch.complex64).contiguous())
        g = g.to(torch.float32).contiguous()
        keep.append(g)
        fields[name] = _native.ptr(g)
    return _native.WcFieldRowsGrad(**fields), keep


def _workspace(rows, layers, device):
    n = int(_native.call("wc_field_rows_workspace_bytes", int(rows), int(layers)))
    return torch.empty(n, dtype=torch.uint8, device=device)


def _opt(t, shape, device):
    if t is None:
        return None
    t = t.detach().to(device=device, dtype=torch.float32).contiguous()
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"tangent / cotangent shape {tuple(t.shape)} does not match {tuple(shape)}")
    return t


# ----------------------------------------------------------------------------- forward operators
def field_free_space(source, target, tx_power, tx_polarization, rx_polarization, *,
                     frequency_hz: float) -> dict[str, torch.Tensor]:
    """E = p_tx,perp / (2 k L) e^{-jkL} per row (reference kernels/fields.py:110)."""
    _check("source", source, torch.float32, 2, (3,))
    count = int(source.shape[0])
    _check_endpoints(count, source, target, tx_power, tx_polarization, rx_polarization, "free-space field")
    _check_positive_frequency(frequency_hz)
    out = _alloc_rows(count, source.device)
    p, c = _native.ptr, _Keep()
    _native.call("wc_field_free_space", count, c(source), c(target), c(tx_power), c(tx_polarization),
                 c(rx_polarization), float(frequency_hz), p(out["field_vector"]), p(out["coefficient"]), p(out["path_field"]),
                 p(out["path_gain"]), p(out["path_length_m"]), p(out["delay_s"]), p(out["direction"]),
                 _native.stream_handle(source.device))
    return out


def field_project_complex3(field_vector, direction, rx_polarization) -> dict[str, torch.Tensor]:
    """coefficient = p_rx,perp . E, path_gain = |coefficient|^2 (reference kernels/fields.py:149)."""
    _check("field_vector", field_vector, torch.complex64, 2)
    _check("direction", direction, torch.float32, 2, (3,))
    _check("rx_polarization", rx_polarization, torch.float32, 2, (3,))
    count = int(field_vector.shape[0])
    if tuple(field_vector.shape) != (count, 3):
        raise ValueError("field_vector must have shape (N, 3)")
    if tuple(direction.shape) != (count, 3) or tuple(rx_polarization.shape) != (count, 3):
        raise ValueError("direction and rx_polarization must match field_vector rows")
    coefficient = torch.empty(count, dtype=torch.complex64, device=field_vector.device)
    path_gain = torch.empty(count, dtype=torch.float32, device=field_vector.device)
    p, c = _native.ptr, _Keep()
    _native.call("wc_field_project_complex3", count, c(field_vector), c(direction), c(rx_polarization),
                 p(coefficient), p(path_gain),
                 _native.stream_handle(field_vector.device))
    return {"coefficient": coefficient, "path_gain": path_gain}


def _check_reflection(source, target, interaction_positions, interaction_normals, tx_power, tx_polarization,
                      rx_polarization, eps_r, sigma_e, mu_r, gain, thickness):
    _check("source", source, torch.float32, 2, (3,))
    count = int(source.shape[0])
    _check_endpoints(count, source, target, tx_power, tx_polarization, rx_polarization, "reflection field")
    for name, value in (("interaction_positions", interaction_positions),
                        ("interaction_normals", interaction_normals)):
        _check(name, value, torch.float32, 3, (3,))
    depth = int(interaction_positions.shape[1])
    if tuple(interaction_positions.shape) != (count, depth, 3) or depth <= 0:
        raise ValueError("interaction_positions must have shape (N, D, 3), D > 0")
    if depth > MAX_REFLECTION_DEPTH:
        raise ValueError(f"reflection depth must be at most {MAX_REFLECTION_DEPTH}")
    if interaction_normals.shape != interaction_positions.shape:
        raise ValueError("interaction_normals must match interaction_positions")
    for name, value in (("eps_r", eps_r), ("sigma_e", sigma_e), ("mu_r", mu_r), ("gain", gain),
                        ("thickness", thickness)):
        _check(name, value, torch.float32, 2)
        if tuple(value.shape) != (count, depth):
            raise ValueError(f"{name} must have shape (N, D)")
    return count, depth


def field_reflection_sequence(source, target, interaction_positions, interaction_normals, tx_power, tx_polarization,
                              rx_polarization, eps_r, sigma_e, mu_r, gain, thickness, *,
                              frequency_hz: float) -> dict[str, torch.Tensor]:
    """Jones transport through D thin-slab reflections (reference kernels/fields.py:180)."""
    count, depth = _check_reflection(source, target, interaction_positions, interaction_normals, tx_power,
                                     tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness)
    _check_positive_frequency(frequency_hz)
    out = _alloc_rows(count, source.device)
    p, c = _native.ptr, _Keep()
    _native.call("wc_field_reflection_sequence", count, depth, c(source), c(target), c(interaction_positions),
                 c(interaction_normals), c(tx_power), c(tx_polarization), c(rx_polarization), c(eps_r), c(sigma_e),
                 c(mu_r), c(gain), c(thickness), float(frequency_hz), p(out["field_vector"]), p(out["coefficient"]),
                 p(out["path_field"]), p(out["path_gain"]), p(out["path_length_m"]), p(out["delay_s"]),
                 p(out["direction"]), _native.stream_handle(source.device))
    return out


def _check_transmission(path_valid, source, target, interaction_positions, interaction_normals,
                        interaction_material_id, interaction_valid, tx_power, tx_polarization, rx_polarization,
                        layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r):
    _check("path_valid", path_valid, torch.bool, 1)
    _check("source", source, torch.float32, 2, (3,))
    count = int(source.shape[0])
    _check_endpoints(count, source, target, tx_power, tx_polarization, rx_polarization, "transmission endpoint")
    for name, value in (("interaction_positions", interaction_positions),
                        ("interaction_normals", interaction_normals)):
        _check(name, value, torch.float32, 3, (3,))
    _check("interaction_material_id", interaction_material_id, torch.int32, 2)
    _check("interaction_valid", interaction_valid, torch.bool, 2)
    if tuple(path_valid.shape) != (count,):
        raise ValueError("path_valid must have shape (N,)")
    depth = int(interaction_positions.shape[1])
    if tuple(interaction_positions.shape) != (count, depth, 3) or depth <= 0:
        raise ValueError("interaction_positions must have shape (N, D, 3), D > 0")
    if depth > MAX_TRANSMISSION_DEPTH:
        raise ValueError(f"transmission depth must be at most {MAX_TRANSMISSION_DEPTH}")
    if interaction_normals.shape != interaction_positions.shape:
        raise ValueError("interaction_normals must match interaction_positions")
    if tuple(interaction_material_id.shape) != (count, depth):
        raise ValueError("interaction_material_id must have shape (N, D)")
    if tuple(interaction_valid.shape) != (count, depth):
        raise ValueError("interaction_valid must have shape (N, D)")
    validate_layer_csr(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                       source.device)
    return count, depth


def _transmission_view(count, depth, frequency_hz, path_valid, source, target, interaction_normals,
                       interaction_material_id, interaction_valid, tx_power, tx_polarization, rx_polarization,
                       tables, keep):
    c = lambda t: keep.append(t.contiguous()) or _native.ptr(keep[-1])  # noqa: E731
    return _native.WcTransmissionRows(
        count=count, depth=depth, frequency_hz=float(frequency_hz), path_valid=c(path_valid),
        interaction_valid=c(interaction_valid), interaction_material_id=c(interaction_material_id),
        source=c(source), target=c(target), interaction_normals=c(interaction_normals), tx_power=c(tx_power),
        tx_polarization=c(tx_polarization), rx_polarization=c(rx_polarization), layers=tables)


def field_transmission_sequence(path_valid, source, target, interaction_positions, interaction_normals,
                                interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                layer_sigma_e, layer_mu_r, *, frequency_hz: float) -> dict[str, torch.Tensor]:
    """Straight thin-sheet transmission chain (reference kernels/fields.py:245):
    per valid interaction the layered t_TE / t_TM on the wall's s/p basis,
    carrier phase over L - sum(thickness cos), amplitude 1 / (2 k L).
    Invalid rows (or a valid interaction naming no material) give zeros."""
    count, depth = _check_transmission(path_valid, source, target, interaction_positions, interaction_normals,
                                       interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                       rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                       layer_sigma_e, layer_mu_r)
    _check_positive_frequency(frequency_hz)
    out = _alloc_rows(count, source.device)
    if count == 0:
        return out
    keep: list = []
    tables = layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                          keep)
    view = _transmission_view(count, depth, frequency_hz, path_valid, source, target, interaction_normals,
                              interaction_material_id, interaction_valid, tx_power, tx_polarization,
                              rx_polarization, tables, keep)
    _native.call("wc_field_transmission_sequence", ctypes.byref(view), ctypes.byref(_rows_out_view(out)),
                 _native.stream_handle(source.device))
    return out


# ----------------------------------------------------------------------------- free space AD
def _grads_dict(grad_field_vector, grad_coefficient, grad_path_field, grad_path_gain, grad_path_length, grad_delay,
                grad_direction=None):
    return {"field_vector": grad_field_vector, "coefficient": grad_coefficient, "path_field": grad_path_field,
            "path_gain": grad_path_gain, "path_length_m": grad_path_length, "delay_s": grad_delay,
            "direction": grad_direction}


def field_free_space_backward(source, target, tx_power, tx_polarization, rx_polarization, *, frequency_hz: float,
                              grad_field_vector=None, grad_coefficient=None, grad_path_field=None,
                              grad_path_gain=None, grad_path_length=None, grad_delay=None, grad_direction=None,
                              need_grad_frequency: bool = True,
                              need_grad_geometry: bool = False) -> dict[str, torch.Tensor | None]:
    """Adjoint of :func:`field_free_space` (reference kernels/fields.py:551):
    ``grad_frequency`` [1] and, with ``need_grad_geometry``, ``grad_source`` /
    ``grad_target`` [N, 3]."""
    count = int(source.shape[0])
    device = source.device
    _check_positive_frequency(frequency_hz)
    view, keep = _grad_view(_grads_dict(grad_field_vector, grad_coefficient, grad_path_field, grad_path_gain,
                                        grad_path_length, grad_delay, grad_direction))
    g_src = torch.empty((count, 3), dtype=torch.float32, device=device) if need_grad_geometry else None
    g_tgt = torch.empty((count, 3), dtype=torch.float32, device=device) if need_grad_geometry else None
    g_freq = torch.empty(1, dtype=torch.float32, device=device)
    ws = _workspace(count, 0, device)
    p, c = _native.ptr, _Keep()
    _native.call("wc_field_free_space_vjp", count, c(source), c(target), c(tx_power), c(tx_polarization),
                 c(rx_polarization), float(frequency_hz), ctypes.byref(view), p(g_src), p(g_tgt), p(g_freq), p(ws),
                 _native.stream_handle(device))
    del keep
    return {"grad_frequency": g_freq if need_grad_frequency else None, "grad_source": g_src, "grad_target": g_tgt}


def field_free_space_jvp(source, target, tx_power, tx_polarization, rx_polarization, *, frequency_hz: float,
                         tangent_frequency: float, tangent_source=None,
                         tangent_target=None) -> dict[str, torch.Tensor]:
    """Tangents of the seven outputs (reference kernels/fields.py:583)."""
    count = int(source.shape[0])
    device = source.device
    _check_positive_frequency(frequency_hz)
    ts = _opt(tangent_source, (count, 3), device)
    tt = _opt(tangent_target, (count, 3), device)
    out = _alloc_rows(count, device)
    p, c = _native.ptr, _Keep()
    _native.call("wc_field_free_space_jvp", count, c(source), c(target), c(tx_power), c(tx_polarization),
                 c(rx_polarization), float(frequency_hz), float(tangent_frequency), p(ts), p(tt),
                 ctypes.byref(_rows_out_view(out)), _native.stream_handle(device))
    return out


# ----------------------------------------------------------------------------- reflection AD
def _reflection_rows(count, depth, frequency_hz, tensors: dict, keep):
    fields = {}
    for name in ("source", "target", "interaction_positions", "interaction_normals", "tx_power", "tx_polarization",
                 "rx_polarization", "eps_r", "sigma_e", "mu_r", "gain", "thickness"):
        t = tensors.get(name)
        if t is None:
            fields[name] = None
        else:
            t = t.detach().to(torch.float32).contiguous()
            keep.append(t)
            fields[name] = _native.ptr(t)
    return _native.WcReflectionRows(count=count, depth=depth, frequency_hz=float(frequency_hz), **fields)


def field_reflection_sequence_backward(source, target, interaction_positions, interaction_normals, tx_power,
                                       tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness, *,
                                       frequency_hz: float, grad_field_vector=None, grad_coefficient=None,
                                       grad_path_field=None, grad_path_gain=None, grad_path_length=None,
                                       grad_delay=None, grad_direction=None, need_grad_eps_r: bool = True,
                                       need_grad_sigma_e: bool = True, need_grad_gain: bool = False,
                                       need_grad_thickness: bool = True, need_grad_frequency: bool = True,
                                       need_grad_geometry: bool = False) -> dict[str, torch.Tensor | None]:
    """Adjoint of :func:`field_reflection_sequence` (reference kernels/fields.py:607)."""
    count, depth = _check_reflection(source, target, interaction_positions, interaction_normals, tx_power,
                                     tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness)
    _check_positive_frequency(frequency_hz)
    device = source.device
    keep: list = []
    rows = _reflection_rows(count, depth, frequency_hz, dict(
        source=source, target=target, interaction_positions=interaction_positions,
        interaction_normals=interaction_normals, tx_power=tx_power, tx_polarization=tx_polarization,
        rx_polarization=rx_polarization, eps_r=eps_r, sigma_e=sigma_e, mu_r=mu_r, gain=gain, thickness=thickness),
        keep)
    view, keep_g = _grad_view(_grads_dict(grad_field_vector, grad_coefficient, grad_path_field, grad_path_gain,
                                          grad_path_length, grad_delay, grad_direction))
    e = lambda need, shape: torch.empty(shape, dtype=torch.float32, device=device) if need else None  # noqa: E731
    out = {
        "grad_eps_r": e(need_grad_eps_r, (count, depth)),
        "grad_sigma_e": e(need_grad_sigma_e, (count, depth)),
        "grad_gain": e(need_grad_gain, (count, depth)),
        "grad_thickness": e(need_grad_thickness, (count, depth)),
        "grad_frequency": torch.empty(1, dtype=torch.float32, device=device),
        "grad_source": e(need_grad_geometry, (count, 3)),
        "grad_target": e(need_grad_geometry, (count, 3)),
        "grad_interaction_positions": e(need_grad_geometry, (count, depth, 3)),
        "grad_interaction_normals": e(need_grad_geometry, (count, depth, 3)),
    }
    p = _native.ptr
    gout = _native.WcReflectionRowsGradOut(**{k: p(v) for k, v in out.items()})
    ws = _workspace(count, 0, device)
    _native.call("wc_field_reflection_sequence_vjp", ctypes.byref(rows), ctypes.byref(view), ctypes.byref(gout),
                 p(ws), _native.stream_handle(device))
    del keep, keep_g
    if not need_grad_frequency:
        out["grad_frequency"] = None
    return out


def field_reflection_sequence_jvp(source, target, interaction_positions, interaction_normals, tx_power,
                                  tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness, *,
                                  frequency_hz: float, tangent_eps_r=None, tangent_sigma_e=None, tangent_gain=None,
                                  tangent_thickness=None, tangent_frequency: float = 0.0, tangent_source=None,
                                  tangent_target=None, tangent_interaction_positions=None,
                                  tangent_interaction_normals=None) -> dict[str, torch.Tensor]:
    """Tangents of the seven outputs (reference kernels/fields.py:665)."""
    count, depth = _check_reflection(source, target, interaction_positions, interaction_normals, tx_power,
                                     tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness)
    _check_positive_frequency(frequency_hz)
    device = source.device
    keep: list = []
    rows = _reflection_rows(count, depth, frequency_hz, dict(
        source=source, target=target, interaction_positions=interaction_positions,
        interaction_normals=interaction_normals, tx_power=tx_power, tx_polarization=tx_polarization,
        rx_polarization=rx_polarization, eps_r=eps_r, sigma_e=sigma_e, mu_r=mu_r, gain=gain, thickness=thickness),
        keep)
    tangents = _reflection_rows(count, depth, frequency_hz, dict(
        source=_opt(tangent_source, (count, 3), device), target=_opt(tangent_target, (count, 3), device),
        interaction_positions=_opt(tangent_interaction_positions, (count, depth, 3), device),
        interaction_normals=_opt(tangent_interaction_normals, (count, depth, 3), device),
        eps_r=_opt(tangent_eps_r, (count, depth), device), sigma_e=_opt(tangent_sigma_e, (count, depth), device),
        gain=_opt(tangent_gain, (count, depth), device), thickness=_opt(tangent_thickness, (count, depth), device)),
        keep)
    out = _alloc_rows(count, device)
    _native.call("wc_field_reflection_sequence_jvp", ctypes.byref(rows), ctypes.byref(tangents),
                 float(tangent_frequency), ctypes.byref(_rows_out_view(out)), _native.stream_handle(device))
    del keep
    return out


# ----------------------------------------------------------------------------- transmission AD
def field_transmission_sequence_backward(path_valid, source, target, interaction_positions, interaction_normals,
                                         interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                         rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                         layer_sigma_e, layer_mu_r, *, frequency_hz: float, grad_field_vector=None,
                                         grad_coefficient=None, grad_path_field=None, grad_path_gain=None,
                                         grad_path_length=None, grad_delay=None,
                                         need_grad_layer_thickness: bool = True, need_grad_layer_eps_r: bool = True,
                                         need_grad_layer_sigma_e: bool = True, need_grad_frequency: bool = True,
                                         need_grad_geometry: bool = False) -> dict[str, torch.Tensor | None]:
    """Adjoint of :func:`field_transmission_sequence` (reference
    kernels/fields.py:710). Layer gradients sum over every row that crosses
    the layer (float64 sums); interaction positions do not enter the field,
    so their gradient is zero."""
    count, depth = _check_transmission(path_valid, source, target, interaction_positions, interaction_normals,
                                       interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                       rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                       layer_sigma_e, layer_mu_r)
    _check_positive_frequency(frequency_hz)
    device = source.device
    layers = int(layer_thickness_m.shape[0])
    e = lambda need, shape: torch.empty(shape, dtype=torch.float32, device=device) if need else None  # noqa: E731
    out = {
        "grad_layer_thickness_m": e(need_grad_layer_thickness, (layers,)),
        "grad_layer_eps_r": e(need_grad_layer_eps_r, (layers,)),
        "grad_layer_sigma_e": e(need_grad_layer_sigma_e, (layers,)),
        "grad_frequency": torch.empty(1, dtype=torch.float32, device=device),
        "grad_source": e(need_grad_geometry, (count, 3)),
        "grad_target": e(need_grad_geometry, (count, 3)),
        "grad_interaction_positions": (torch.zeros((count, depth, 3), dtype=torch.float32, device=device)
                                       if need_grad_geometry else None),
        "grad_interaction_normals": e(need_grad_geometry, (count, depth, 3)),
    }
    keep: list = []
    tables = layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                          keep)
    view = _transmission_view(count, depth, frequency_hz, path_valid, source, target, interaction_normals,
                              interaction_material_id, interaction_valid, tx_power, tx_polarization,
                              rx_polarization, tables, keep)
    gview, keep_g = _grad_view(_grads_dict(grad_field_vector, grad_coefficient, grad_path_field, grad_path_gain,
                                           grad_path_length, grad_delay))
    ws = _workspace(count, layers, device)
    p = _native.ptr
    if layers == 0:
        for k in ("grad_layer_thickness_m", "grad_layer_eps_r", "grad_layer_sigma_e"):
            if out[k] is not None:
                out[k] = torch.zeros(0, dtype=torch.float32, device=device)
    _native.call("wc_field_transmission_sequence_vjp", ctypes.byref(view), ctypes.byref(gview),
                 p(out["grad_source"]), p(out["grad_target"]), p(out["grad_interaction_normals"]),
                 p(out["grad_layer_thickness_m"]), p(out["grad_layer_eps_r"]), p(out["grad_layer_sigma_e"]),
                 p(out["grad_frequency"]), p(ws), _native.stream_handle(device))
    del keep, keep_g
    if not need_grad_frequency:
        out["grad_frequency"] = None
    return out


def field_transmission_sequence_jvp(path_valid, source, target, interaction_positions, interaction_normals,
                                    interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                    rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                    layer_sigma_e, layer_mu_r, *, frequency_hz: float,
                                    tangent_layer_thickness_m=None, tangent_layer_eps_r=None,
                                    tangent_layer_sigma_e=None, tangent_frequency: float = 0.0, tangent_source=None,
                                    tangent_target=None, tangent_interaction_positions=None,
                                    tangent_interaction_normals=None) -> dict[str, torch.Tensor]:
    """Tangents of the six transmission outputs (reference kernels/fields.py:771;
    no direction tangent, as in the reference)."""
    count, depth = _check_transmission(path_valid, source, target, interaction_positions, interaction_normals,
                                       interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                       rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                       layer_sigma_e, layer_mu_r)
    _check_positive_frequency(frequency_hz)
    device = source.device
    layers = int(layer_thickness_m.shape[0])
    del tangent_interaction_positions  # positions do not enter the field
    keep: list = []
    tables = layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                          keep)
    view = _transmission_view(count, depth, frequency_hz, path_valid, source, target, interaction_normals,
                              interaction_material_id, interaction_valid, tx_power, tx_polarization,
                              rx_polarization, tables, keep)
    tan = [_opt(tangent_source, (count, 3), device), _opt(tangent_target, (count, 3), device),
           _opt(tangent_interaction_normals, (count, depth, 3), device),
           _opt(tangent_layer_thickness_m, (layers,), device), _opt(tangent_layer_eps_r, (layers,), device),
           _opt(tangent_layer_sigma_e, (layers,), device)]
    out = _alloc_rows(count, device, with_direction=False)
    if count:
        p = _native.ptr
        _native.call("wc_field_transmission_sequence_jvp", ctypes.byref(view), *[p(t) for t in tan],
                     float(tangent_frequency), ctypes.byref(_rows_out_view(out)), _native.stream_handle(device))
    del keep
    return out


# ----------------------------------------------------------------------------- differentiable forms
def _frequency_value(frequency) -> float:
    if isinstance(frequency, torch.Tensor):
        return float(frequency.detach().reshape(-1)[0])
    return float(frequency)


def _tangent_frequency(t) -> float:
    return 0.0 if t is None else float(t.detach().reshape(-1)[0])


def _freq_grad(g, frequency):
    if not isinstance(frequency, torch.Tensor) or g is None:
        return None
    return g.to(frequency.dtype).reshape(frequency.shape).to(frequency.device)


def _live(*tangents) -> bool:
    """True when any tangent is present and nonzero (autograd materializes
    zero tangents for inputs that carry none)."""
    return any(t is not None and bool(t.ne(0).any()) for t in tangents)


def _cotangents(grads, names):
    return {name: g for name, g in zip(names, grads)}


class _FieldFreeSpaceAdFunction(torch.autograd.Function):
    """Differentiable free space: frequency and both endpoints (reference
    kernels/fields.py:1087); power and polarizations fixed."""

    @staticmethod
    def forward(ctx, source, target, tx_power, tx_polarization, rx_polarization, frequency,  # type: ignore[override]
                frequency_value):
        out = field_free_space(source.detach(), target.detach(), tx_power.detach(), tx_polarization.detach(),
                               rx_polarization.detach(), frequency_hz=frequency_value)
        ctx.inputs = (source.detach(), target.detach(), tx_power.detach(), tx_polarization.detach(),
                      rx_polarization.detach())
        ctx.frequency = frequency
        ctx.frequency_value = frequency_value
        return tuple(out[name] for name in FIELD_OUTPUT_FIELDS)

    @staticmethod
    def backward(ctx, *grads):  # type: ignore[override]
        g = _cotangents(grads, FIELD_OUTPUT_FIELDS)
        need_geo = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        out = field_free_space_backward(*ctx.inputs, frequency_hz=ctx.frequency_value,
                                        grad_field_vector=g["field_vector"], grad_coefficient=g["coefficient"],
                                        grad_path_field=g["path_field"], grad_path_gain=g["path_gain"],
                                        grad_path_length=g["path_length_m"], grad_delay=g["delay_s"],
                                        grad_direction=g["direction"], need_grad_geometry=need_geo)
        return (out["grad_source"], out["grad_target"], None, None, None,
                _freq_grad(out["grad_frequency"], ctx.frequency) if ctx.needs_input_grad[5] else None, None)

    @staticmethod
    def jvp(ctx, t_source, t_target, t_power, t_txp, t_rxp, t_frequency, _):  # type: ignore[override]
        if _live(t_power, t_txp, t_rxp):
            raise ValueError("tx_power and polarizations are fixed inputs of field_free_space_ad")
        out = field_free_space_jvp(*ctx.inputs, frequency_hz=ctx.frequency_value,
                                   tangent_frequency=_tangent_frequency(t_frequency), tangent_source=t_source,
                                   tangent_target=t_target)
        return tuple(out[name] for name in FIELD_OUTPUT_FIELDS)


def field_free_space_ad(source, target, tx_power, tx_polarization, rx_polarization, *, frequency,
                        frequency_value: float | None = None, direction_live: bool = False) -> dict[str, torch.Tensor]:
    """Differentiable :func:`field_free_space` (reference kernels/fields.py:1255).
    ``direction_live`` is accepted for signature parity; the direction output
    always carries its derivative here."""
    del direction_live
    fv = _frequency_value(frequency) if frequency_value is None else float(frequency_value)
    values = _FieldFreeSpaceAdFunction.apply(source, target, tx_power, tx_polarization, rx_polarization, frequency,
                                             fv)
    return dict(zip(FIELD_OUTPUT_FIELDS, values))


class _FieldReflectionSequenceAdFunction(torch.autograd.Function):
    """Differentiable reflection sequence: frequency, hit geometry and every
    per-bounce material scalar except mu_r (reference kernels/fields.py:1283)."""

    @staticmethod
    def forward(ctx, source, target, positions, normals, tx_power, tx_polarization, rx_polarization,  # type: ignore[override]
                eps_r, sigma_e, mu_r, gain, thickness, frequency, frequency_value):
        ins = tuple(t.detach() for t in (source, target, positions, normals, tx_power, tx_polarization,
                                          rx_polarization, eps_r, sigma_e, mu_r, gain, thickness))
        out = field_reflection_sequence(*ins, frequency_hz=frequency_value)
        ctx.inputs = ins
        ctx.frequency = frequency
        ctx.frequency_value = frequency_value
        return tuple(out[name] for name in FIELD_OUTPUT_FIELDS)

    @staticmethod
    def backward(ctx, *grads):  # type: ignore[override]
        g = _cotangents(grads, FIELD_OUTPUT_FIELDS)
        n = ctx.needs_input_grad
        out = field_reflection_sequence_backward(
            *ctx.inputs, frequency_hz=ctx.frequency_value, grad_field_vector=g["field_vector"],
            grad_coefficient=g["coefficient"], grad_path_field=g["path_field"], grad_path_gain=g["path_gain"],
            grad_path_length=g["path_length_m"], grad_delay=g["delay_s"], grad_direction=g["direction"],
            need_grad_eps_r=n[7], need_grad_sigma_e=n[8], need_grad_gain=n[10], need_grad_thickness=n[11],
            need_grad_frequency=n[12], need_grad_geometry=any(n[:4]))
        return (out["grad_source"] if n[0] else None, out["grad_target"] if n[1] else None,
                out["grad_interaction_positions"] if n[2] else None,
                out["grad_interaction_normals"] if n[3] else None, None, None, None, out["grad_eps_r"],
                out["grad_sigma_e"], None, out["grad_gain"], out["grad_thickness"],
                _freq_grad(out["grad_frequency"], ctx.frequency) if n[12] else None, None)

    @staticmethod
    def jvp(ctx, t_src, t_tgt, t_pos, t_nrm, t_power, t_txp, t_rxp, t_eps, t_sig, t_mu, t_gain, t_thick,  # type: ignore[override]
            t_freq, _):
        if _live(t_power, t_txp, t_rxp, t_mu):
            raise ValueError("tx_power, polarizations and mu_r are fixed inputs of field_reflection_sequence_ad")
        out = field_reflection_sequence_jvp(
            *ctx.inputs, frequency_hz=ctx.frequency_value, tangent_eps_r=t_eps, tangent_sigma_e=t_sig,
            tangent_gain=t_gain, tangent_thickness=t_thick, tangent_frequency=_tangent_frequency(t_freq),
            tangent_source=t_src, tangent_target=t_tgt, tangent_interaction_positions=t_pos,
            tangent_interaction_normals=t_nrm)
        return tuple(out[name] for name in FIELD_OUTPUT_FIELDS)


def field_reflection_sequence_ad(source, target, interaction_positions, interaction_normals, tx_power,
                                 tx_polarization, rx_polarization, eps_r, sigma_e, mu_r, gain, thickness, *,
                                 frequency, frequency_value: float | None = None,
                                 direction_live: bool = False) -> dict[str, torch.Tensor]:
    """Differentiable :func:`field_reflection_sequence` (reference kernels/fields.py:1522)."""
    del direction_live
    fv = _frequency_value(frequency) if frequency_value is None else float(frequency_value)
    values = _FieldReflectionSequenceAdFunction.apply(source, target, interaction_positions, interaction_normals,
                                                      tx_power, tx_polarization, rx_polarization, eps_r, sigma_e,
                                                      mu_r, gain, thickness, frequency, fv)
    return dict(zip(FIELD_OUTPUT_FIELDS, values))


class _FieldTransmissionSequenceAdFunction(torch.autograd.Function):
    """Differentiable transmission: frequency, endpoints, wall normals and the
    layer tables' thickness / eps_r / sigma_e (reference kernels/fields.py:1560).
    The direction output normalize(target - source) depends on the endpoints
    only; its derivative (I - d d^T) / L is applied here on both passes."""

    @staticmethod
    def forward(ctx, path_valid, source, target, positions, normals, material_id, interaction_valid,  # type: ignore[override]
                tx_power, tx_polarization, rx_polarization, layer_offset, layer_count, layer_thickness_m,
                layer_eps_r, layer_sigma_e, layer_mu_r, frequency, frequency_value):
        ins = tuple(t.detach() for t in (path_valid, source, target, positions, normals, material_id,
                                          interaction_valid, tx_power, tx_polarization, rx_polarization,
                                          layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                          layer_sigma_e, layer_mu_r))
        out = field_transmission_sequence(*ins, frequency_hz=frequency_value)
        ctx.inputs = ins
        ctx.frequency = frequency
        ctx.frequency_value = frequency_value
        ctx.direction = out["direction"]
        ctx.length = out["path_length_m"]
        return tuple(out[name] for name in FIELD_OUTPUT_FIELDS)

    @staticmethod
    def _direction_map(ctx, v):
        """(I - d d^T) v / L per row (zero on rows with no direction)."""
        d = ctx.direction
        inv = torch.where(ctx.length > 0, 1.0 / ctx.length.clamp_min(1e-30), torch.zeros_like(ctx.length))
        return (v - d * (d * v).sum(-1, keepdim=True)) * inv[:, None]

    @staticmethod
    def backward(ctx, *grads):  # type: ignore[override]
        g = _cotangents(grads, FIELD_OUTPUT_FIELDS)
        n = ctx.needs_input_grad
        out = field_transmission_sequence_backward(
            *ctx.inputs, frequency_hz=ctx.frequency_value, grad_field_vector=g["field_vector"],
            grad_coefficient=g["coefficient"], grad_path_field=g["path_field"], grad_path_gain=g["path_gain"],
            grad_path_length=g["path_length_m"], grad_delay=g["delay_s"], need_grad_layer_thickness=n[12],
            need_grad_layer_eps_r=n[13], need_grad_layer_sigma_e=n[14], need_grad_frequency=n[16],
            need_grad_geometry=any(n[1:5]))
        g_dir = g["direction"]
        if g_dir is not None and out["grad_source"] is not None:
            m = _FieldTransmissionSequenceAdFunction._direction_map(ctx, g_dir.to(torch.float32))
            out["grad_source"] = out["grad_source"] - m
            out["grad_target"] = out["grad_target"] + m
        return (None, out["grad_source"] if n[1] else None, out["grad_target"] if n[2] else None,
                out["grad_interaction_positions"] if n[3] else None,
                out["grad_interaction_normals"] if n[4] else None, None, None, None, None, None, None, None,
                out["grad_layer_thickness_m"], out["grad_layer_eps_r"], out["grad_layer_sigma_e"], None,
                _freq_grad(out["grad_frequency"], ctx.frequency) if n[16] else None, None)

    @staticmethod
    def jvp(ctx, t_valid, t_src, t_tgt, t_pos, t_nrm, t_mat, t_iv, t_power, t_txp, t_rxp, t_off, t_cnt,  # type: ignore[override]
            t_thick, t_eps, t_sig, t_mu, t_freq, _):
        if _live(t_power, t_txp, t_rxp, t_mu):
            raise ValueError("tx_power, polarizations and layer_mu_r are fixed inputs of "
                             "field_transmission_sequence_ad")
        out = field_transmission_sequence_jvp(
            *ctx.inputs, frequency_hz=ctx.frequency_value, tangent_layer_thickness_m=t_thick,
            tangent_layer_eps_r=t_eps, tangent_layer_sigma_e=t_sig, tangent_frequency=_tangent_frequency(t_freq),
            tangent_source=t_src, tangent_target=t_tgt, tangent_interaction_positions=t_pos,
            tangent_interaction_normals=t_nrm)
        rel = torch.zeros_like(ctx.direction)
        if t_tgt is not None:
            rel = rel + t_tgt
        if t_src is not None:
            rel = rel - t_src
        t_dir = _FieldTransmissionSequenceAdFunction._direction_map(ctx, rel)
        return tuple(out[name] for name in FIELD_TANGENT_FIELDS) + (t_dir,)


def field_transmission_sequence_ad(path_valid, source, target, interaction_positions, interaction_normals,
                                   interaction_material_id, interaction_valid, tx_power, tx_polarization,
                                   rx_polarization, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                                   layer_sigma_e, layer_mu_r, *, frequency,
                                   frequency_value: float | None = None) -> dict[str, torch.Tensor]:
    """Differentiable :func:`field_transmission_sequence` (reference kernels/fields.py:1786)."""
    fv = _frequency_value(frequency) if frequency_value is None else float(frequency_value)
    values = _FieldTransmissionSequenceAdFunction.apply(
        path_valid, source, target, interaction_positions, interaction_normals, interaction_material_id,
        interaction_valid, tx_power, tx_polarization, rx_polarization, layer_offset, layer_count, layer_thickness_m,
        layer_eps_r, layer_sigma_e, layer_mu_r, frequency, fv)
    return dict(zip(FIELD_OUTPUT_FIELDS, values))


__all__ = ["field_free_space", "field_free_space_ad", "field_free_space_backward", "field_free_space_jvp",
           "field_project_complex3", "field_reflection_sequence", "field_reflection_sequence_ad",
           "field_reflection_sequence_backward", "field_reflection_sequence_jvp", "field_transmission_sequence",
           "field_transmission_sequence_ad", "field_transmission_sequence_backward",
           "field_transmission_sequence_jvp", "layer_tables", "validate_layer_csr"]
