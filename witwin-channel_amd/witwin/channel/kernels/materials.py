"""Layer-stack material operators (reference witwin/channel/kernels/
materials.py:136-461): ``em_layer_stack_eval`` (amplitude r / t per
polarization and power coefficients R = |r|^2, T = |t|^2 of a layered wall
between vacuum half-spaces, backward Rouard recursion), its adjoint
``em_layer_stack_backward``, tangent ``em_layer_stack_jvp`` and the
differentiable ``em_layer_stack_ad``. Same argument meaning, output keys and
shapes as the reference; the kernels are csrc/wc_rows.hip (stack_*_kernel).
A zero-layer material is transparent (r = 0, t = 1); a material id outside
the tables gives zeros."""

from __future__ import annotations

import ctypes

import torch

from witwin.channel import _native
from witwin.channel.kernels.fields import _check, _check_positive_frequency, _live, layer_tables, validate_layer_csr

EM_LAYER_STACK_FIELDS = _native.LAYER_STACK_FIELDS
_EM_LAYER_STACK_FIELDS = EM_LAYER_STACK_FIELDS


def _check_rows(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e,
                layer_mu_r, frequency_hz):
    _check("cos_theta", cos_theta, torch.float32, 1)
    _check("material_id", material_id, torch.int32, 1)
    count = int(cos_theta.shape[0])
    if tuple(material_id.shape) != (count,):
        raise ValueError("material_id must match cos_theta length")
    if material_id.device != cos_theta.device:
        raise ValueError("material_id must share cos_theta device")
    validate_layer_csr(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                       cos_theta.device)
    _check_positive_frequency(frequency_hz)
    return count


def _alloc(count, device):
    return {name: torch.empty(count, dtype=torch.float32, device=device) for name in EM_LAYER_STACK_FIELDS}


def _view(tensors: dict) -> _native.WcLayerStackOut:
    return _native.WcLayerStackOut(**{name: _native.ptr(tensors.get(name)) for name in EM_LAYER_STACK_FIELDS})


def em_layer_stack_eval(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                        layer_sigma_e, layer_mu_r, *, frequency_hz: float) -> dict[str, torch.Tensor]:
    """Twelve [N] float32 outputs r_te_real .. cap_T_tm (reference kernels/materials.py:151)."""
    count = _check_rows(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                        layer_sigma_e, layer_mu_r, frequency_hz)
    device = cos_theta.device
    out = _alloc(count, device)
    if count == 0:
        return out
    keep: list = []
    tables = layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                          keep)
    cos_c, mat_c = cos_theta.contiguous(), material_id.contiguous()
    _native.call("wc_em_layer_stack_eval", count, _native.ptr(cos_c), _native.ptr(mat_c), ctypes.byref(tables),
                 float(frequency_hz), ctypes.byref(_view(out)), _native.stream_handle(device))
    return out


def em_layer_stack_backward(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                            layer_sigma_e, layer_mu_r, grad_outputs, *, frequency_hz: float, need_cos_theta: bool,
                            need_layers: bool, need_frequency: bool) -> dict[str, torch.Tensor]:
    """Adjoint (reference kernels/materials.py:199): one cotangent slot per
    output (None = zero). Layer gradients sum over every row of the material
    (float64 sums); grad_frequency has shape [1]."""
    if len(grad_outputs) != len(EM_LAYER_STACK_FIELDS):
        raise ValueError("grad_outputs must carry one cotangent slot per stack output")
    count = _check_rows(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                        layer_sigma_e, layer_mu_r, frequency_hz)
    device = cos_theta.device
    layers = int(layer_thickness_m.shape[0])
    grads = {}
    for name, g in zip(EM_LAYER_STACK_FIELDS, grad_outputs):
        if g is not None:
            g = g.detach().to(torch.float32).contiguous()
            if tuple(g.shape) != (count,):
                raise ValueError(f"cotangent of {name} must have shape ({count},)")
        grads[name] = g
    out = {
        "grad_cos_theta": torch.empty(count, dtype=torch.float32, device=device),
        "grad_layer_thickness_m": torch.zeros(layers, dtype=torch.float32, device=device),
        "grad_layer_eps_r": torch.zeros(layers, dtype=torch.float32, device=device),
        "grad_layer_sigma_e": torch.zeros(layers, dtype=torch.float32, device=device),
        "grad_frequency": torch.zeros(1, dtype=torch.float32, device=device),
    }
    keep: list = []
    tables = layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                          keep)
    ws = torch.empty(int(_native.call("wc_field_rows_workspace_bytes", count, layers)), dtype=torch.uint8,
                     device=device)
    p = _native.ptr
    cos_c, mat_c = cos_theta.contiguous(), material_id.contiguous()
    _native.call("wc_em_layer_stack_vjp", count, p(cos_c), p(mat_c), ctypes.byref(tables), float(frequency_hz),
                 ctypes.byref(_view(grads)), p(out["grad_cos_theta"]),
                 p(out["grad_layer_thickness_m"]) if layers else None,
                 p(out["grad_layer_eps_r"]) if layers else None, p(out["grad_layer_sigma_e"]) if layers else None,
                 p(out["grad_frequency"]), p(ws), _native.stream_handle(device))
    if not need_cos_theta:
        out["grad_cos_theta"] = None
    if not need_layers:
        out["grad_layer_thickness_m"] = out["grad_layer_eps_r"] = out["grad_layer_sigma_e"] = None
    if not need_frequency:
        out["grad_frequency"] = None
    return out


def em_layer_stack_jvp(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                       layer_sigma_e, layer_mu_r, *, frequency_hz: float, tangent_cos_theta, tangent_layer_thickness,
                       tangent_layer_eps_r, tangent_layer_sigma_e, tangent_frequency: float) -> dict[str, torch.Tensor]:
    """Tangents of the twelve outputs (reference kernels/materials.py:235); None tangents are zero."""
    count = _check_rows(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                        layer_sigma_e, layer_mu_r, frequency_hz)
    device = cos_theta.device
    layers = int(layer_thickness_m.shape[0])
    out = _alloc(count, device)
    if count == 0:
        return out

    def tan(t, n):
        if t is None:
            return None
        t = t.detach().to(device=device, dtype=torch.float32).contiguous()
        if t.numel() != n:
            raise ValueError("tangent shape does not match its primal")
        return t

    tangents = [tan(tangent_cos_theta, count), tan(tangent_layer_thickness, layers),
                tan(tangent_layer_eps_r, layers), tan(tangent_layer_sigma_e, layers)]
    keep: list = []
    tables = layer_tables(layer_offset, layer_count, layer_thickness_m, layer_eps_r, layer_sigma_e, layer_mu_r,
                          keep)
    p = _native.ptr
    cos_c, mat_c = cos_theta.contiguous(), material_id.contiguous()
    _native.call("wc_em_layer_stack_jvp", count, p(cos_c), p(mat_c), ctypes.byref(tables), float(frequency_hz),
                 *[p(t) for t in tangents], float(tangent_frequency), ctypes.byref(_view(out)),
                 _native.stream_handle(device))
    return out


class _EmLayerStackAdFunction(torch.autograd.Function):
    """Differentiable layer stack: cos_theta, the CSR thickness / eps_r /
    sigma_e and the frequency (reference kernels/materials.py:293); mu_r and
    the CSR topology are fixed."""

    @staticmethod
    def forward(ctx, cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,  # type: ignore[override]
                layer_sigma_e, layer_mu_r, frequency, frequency_value):
        ins = tuple(t.detach() for t in (cos_theta, material_id, layer_offset, layer_count, layer_thickness_m,
                                          layer_eps_r, layer_sigma_e, layer_mu_r))
        out = em_layer_stack_eval(*ins, frequency_hz=frequency_value)
        ctx.inputs = ins
        ctx.frequency = frequency
        ctx.frequency_value = frequency_value
        return tuple(out[name] for name in EM_LAYER_STACK_FIELDS)

    @staticmethod
    def backward(ctx, *grads):  # type: ignore[override]
        n = ctx.needs_input_grad
        if n[7]:
            raise ValueError("em_layer_stack_ad: layer_mu_r is a fixed input")
        out = em_layer_stack_backward(*ctx.inputs, grads, frequency_hz=ctx.frequency_value, need_cos_theta=n[0],
                                      need_layers=any(n[4:7]), need_frequency=n[8])
        g_freq = None
        if n[8] and isinstance(ctx.frequency, torch.Tensor):
            g_freq = out["grad_frequency"].to(ctx.frequency.dtype).reshape(ctx.frequency.shape)
        return (out["grad_cos_theta"], None, None, None, out["grad_layer_thickness_m"] if n[4] else None,
                out["grad_layer_eps_r"] if n[5] else None, out["grad_layer_sigma_e"] if n[6] else None, None, g_freq,
                None)

    @staticmethod
    def jvp(ctx, t_cos, t_mat, t_off, t_cnt, t_thick, t_eps, t_sig, t_mu, t_freq, _):  # type: ignore[override]
        if _live(t_mu):
            raise ValueError("em_layer_stack_ad: layer_mu_r is a fixed input")
        out = em_layer_stack_jvp(*ctx.inputs, frequency_hz=ctx.frequency_value, tangent_cos_theta=t_cos,
                                 tangent_layer_thickness=t_thick, tangent_layer_eps_r=t_eps,
                                 tangent_layer_sigma_e=t_sig,
                                 tangent_frequency=0.0 if t_freq is None else float(t_freq.reshape(-1)[0]))
        return tuple(out[name] for name in EM_LAYER_STACK_FIELDS)


def em_layer_stack_ad(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m, layer_eps_r,
                      layer_sigma_e, layer_mu_r, *, frequency,
                      frequency_value: float | None = None) -> dict[str, torch.Tensor]:
    """Differentiable :func:`em_layer_stack_eval` (reference kernels/materials.py:433)."""
    if frequency_value is None:
        frequency_value = float(frequency.detach().reshape(-1)[0]) if isinstance(frequency, torch.Tensor) \
            else float(frequency)
    values = _EmLayerStackAdFunction.apply(cos_theta, material_id, layer_offset, layer_count, layer_thickness_m,
                                           layer_eps_r, layer_sigma_e, layer_mu_r, frequency, float(frequency_value))
    return dict(zip(EM_LAYER_STACK_FIELDS, values))


__all__ = ["EM_LAYER_STACK_FIELDS", "em_layer_stack_ad", "em_layer_stack_backward", "em_layer_stack_eval",
           "em_layer_stack_jvp", "validate_layer_csr"]
