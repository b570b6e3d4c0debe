"""Flat accumulation facades (reference witwin/channel/kernels/
deterministic.py:23-390): ``deterministic_accumulate_flat`` (float32 rows),
its float64 companion, the adjoint ``deterministic_accumulate_flat_backward``,
the tangent ``deterministic_accumulate_flat_jvp`` and the differentiable
``deterministic_accumulate_flat_ad`` (reverse and forward mode), all on
``wc_accumulate_rows*`` (csrc/wc_accum.hip). Same argument meaning, output
keys and shapes as the reference: ``power_total`` / ``field_total_real`` /
``field_total_imag`` [T, R] and ``component_power`` /
``component_field_real`` / ``component_field_imag`` [6, T, R] (slots los,
reflection, diffraction, transmission, scattering, coupled)."""

from __future__ import annotations

import ctypes

import torch

from witwin.channel import _native

ACCUM_FIELDS = ("power_total", "field_total_real", "field_total_imag", "component_power", "component_field_real",
                "component_field_imag")
SLOTS = 6


def _check_rows(valid, tx_id, rx_id, component_id, values, dtype, num_tx, num_rx, scattering_combine_domain):
    if int(scattering_combine_domain) not in (0, 1):
        raise ValueError("scattering_combine_domain must be 0 (power) or 1 (coherent)")
    if num_tx < 0 or num_rx < 0:
        raise ValueError("num_tx and num_rx must be non-negative")
    for name, t, want in (("valid", valid, torch.bool), ("tx_id", tx_id, torch.int32), ("rx_id", rx_id, torch.int32),
                          ("component_id", component_id, torch.int32)):
        if not isinstance(t, torch.Tensor) or t.dtype != want or t.dim() != 1:
            raise ValueError(f"{name} must be a 1-D {want} tensor")
        if t.shape != tx_id.shape:
            raise ValueError(f"{name} must match tx_id shape")
    for name, t in values.items():
        if t is None:
            continue
        if t.dtype != dtype or t.dim() != 1 or t.shape != tx_id.shape:
            raise ValueError(f"{name} must be a 1-D {dtype} tensor matching tx_id shape")


def _rows_view(valid, tx_id, rx_id, component_id, gain, real, imag, num_tx, num_rx, coherent, scattering, dtype):
    p = _native.ptr
    return _native.WcAccumRows(count=int(tx_id.shape[0]), num_tx=int(num_tx), num_rx=int(num_rx),
                               coherent=1 if coherent else 0, scattering_coherent=int(scattering),
                               dtype=1 if dtype == torch.float64 else 0, valid=p(valid.contiguous()),
                               tx_id=p(tx_id), rx_id=p(rx_id), component_id=p(component_id), path_gain=p(gain),
                               field_real=p(real), field_imag=p(imag))


def _alloc(num_tx, num_rx, dtype, device) -> dict[str, torch.Tensor]:
    t, r = int(num_tx), int(num_rx)
    out = {name: torch.empty((t, r), dtype=dtype, device=device) for name in ACCUM_FIELDS[:3]}
    out.update({name: torch.empty((SLOTS, t, r), dtype=dtype, device=device) for name in ACCUM_FIELDS[3:]})
    return out


def _out_view(out: dict[str, torch.Tensor]):
    return _native.WcAccumOut(**{name: _native.ptr(out[name]) for name in ACCUM_FIELDS})


def _workspace(count, num_tx, num_rx, device):
    return torch.empty(int(_native.call("wc_accumulate_rows_workspace_bytes", int(count), int(num_tx), int(num_rx))),
                       dtype=torch.uint8, device=device)


def _accumulate(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, num_tx, num_rx, coherent,
                scattering_combine_domain, dtype):
    _check_rows(valid, tx_id, rx_id, component_id,
                {"path_gain": path_gain, "field_real": field_real, "field_imag": field_imag}, dtype, num_tx, num_rx,
                scattering_combine_domain)
    device = tx_id.device
    gain, real, imag = (x.contiguous() for x in (path_gain, field_real, field_imag))
    out = _alloc(num_tx, num_rx, dtype, device)
    view = _rows_view(valid, tx_id, rx_id, component_id, gain, real, imag, num_tx, num_rx, coherent,
                      scattering_combine_domain, dtype)
    ws = _workspace(tx_id.shape[0], num_tx, num_rx, device)
    _native.call("wc_accumulate_rows", ctypes.byref(view), ctypes.byref(_out_view(out)), _native.ptr(ws),
                 _native.stream_handle(device))
    return out


def deterministic_accumulate_flat(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, *,
                                  num_tx: int, num_rx: int, coherent: bool,
                                  scattering_combine_domain: int = 0) -> dict[str, torch.Tensor]:
    """float32 rows -> the six accumulated outputs (reference
    kernels/deterministic.py:23-102)."""
    return _accumulate(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, num_tx, num_rx,
                       coherent, scattering_combine_domain, torch.float32)


def deterministic_accumulate_flat_fwd64(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, *,
                                        num_tx: int, num_rx: int, coherent: bool,
                                        scattering_combine_domain: int = 0) -> dict[str, torch.Tensor]:
    """The same accumulation for float64 rows (the reference's gradcheck
    companion)."""
    return _accumulate(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, num_tx, num_rx,
                       coherent, scattering_combine_domain, torch.float64)


def deterministic_accumulate_flat_backward(valid, tx_id, rx_id, component_id, component_field_real,
                                           component_field_imag, field_total_real, field_total_imag, power_total, *,
                                           grad_power_total=None, grad_field_total_real=None,
                                           grad_field_total_imag=None, grad_component_power=None,
                                           grad_component_field_real=None, grad_component_field_imag=None,
                                           num_tx: int, num_rx: int, coherent: bool,
                                           scattering_combine_domain: int = 0) -> dict[str, torch.Tensor]:
    """Per-row gradients of path_gain / field_real / field_imag from the
    output cotangents (reference kernels/deterministic.py:115-153)."""
    dtype = power_total.dtype
    _check_rows(valid, tx_id, rx_id, component_id, {}, dtype, num_tx, num_rx, scattering_combine_domain)
    device = tx_id.device
    c = lambda x: None if x is None else x.to(dtype).contiguous()  # noqa: E731
    fwd = {"power_total": c(power_total), "field_total_real": c(field_total_real),
           "field_total_imag": c(field_total_imag), "component_power": None,
           "component_field_real": c(component_field_real), "component_field_imag": c(component_field_imag)}
    grads = {"power_total": c(grad_power_total), "field_total_real": c(grad_field_total_real),
             "field_total_imag": c(grad_field_total_imag), "component_power": c(grad_component_power),
             "component_field_real": c(grad_component_field_real),
             "component_field_imag": c(grad_component_field_imag)}
    count = int(tx_id.shape[0])
    out = {name: torch.empty(count, dtype=dtype, device=device)
           for name in ("grad_path_gain", "grad_field_real", "grad_field_imag")}
    view = _rows_view(valid, tx_id, rx_id, component_id, None, None, None, num_tx, num_rx, coherent,
                      scattering_combine_domain, dtype)
    p = _native.ptr
    _native.call("wc_accumulate_rows_backward", ctypes.byref(view),
                 ctypes.byref(_native.WcAccumOut(**{k: p(v) for k, v in fwd.items()})),
                 ctypes.byref(_native.WcAccumOut(**{k: p(v) for k, v in grads.items()})),
                 p(out["grad_path_gain"]), p(out["grad_field_real"]), p(out["grad_field_imag"]),
                 _native.stream_handle(device))
    return out


def deterministic_accumulate_flat_jvp(valid, tx_id, rx_id, component_id, component_field_real, component_field_imag,
                                      power_total, *, tangent_path_gain=None, tangent_field_real=None,
                                      tangent_field_imag=None, num_tx: int, num_rx: int, coherent: bool,
                                      scattering_combine_domain: int = 0) -> dict[str, torch.Tensor]:
    """Tangents of the six outputs from row tangents (reference
    kernels/deterministic.py:156-183)."""
    dtype = power_total.dtype
    _check_rows(valid, tx_id, rx_id, component_id,
                {"tangent_path_gain": tangent_path_gain, "tangent_field_real": tangent_field_real,
                 "tangent_field_imag": tangent_field_imag}, dtype, num_tx, num_rx, scattering_combine_domain)
    device = tx_id.device
    c = lambda x: None if x is None else x.contiguous()  # noqa: E731
    out = _alloc(num_tx, num_rx, dtype, device)
    view = _rows_view(valid, tx_id, rx_id, component_id, c(tangent_path_gain), c(tangent_field_real),
                      c(tangent_field_imag), num_tx, num_rx, coherent, scattering_combine_domain, dtype)
    p = _native.ptr
    fwd = _native.WcAccumOut(power_total=p(power_total.contiguous()), field_total_real=None, field_total_imag=None,
                             component_power=None, component_field_real=p(component_field_real.contiguous()),
                             component_field_imag=p(component_field_imag.contiguous()))
    ws = _workspace(tx_id.shape[0], num_tx, num_rx, device)
    _native.call("wc_accumulate_rows_jvp", ctypes.byref(view), ctypes.byref(fwd), ctypes.byref(_out_view(out)),
                 p(ws), _native.stream_handle(device))
    return out


class _DeterministicAccumulateFlatAdFunction(torch.autograd.Function):
    """Differentiable flat accumulation: native forward, native adjoint
    (backward) and native tangent (jvp), with the |.|^2 and sqrt cell
    nonlinearities linearized at the saved forward outputs."""

    @staticmethod
    def forward(ctx, path_gain, field_real, field_imag, valid, tx_id, rx_id, component_id, num_tx, num_rx,  # type: ignore[override]
                coherent, scattering):
        out = _accumulate(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, num_tx, num_rx,
                          coherent, scattering, path_gain.dtype)
        ctx.gates = (valid, tx_id, rx_id, component_id, num_tx, num_rx, coherent, scattering)
        ctx.fwd = out
        return tuple(out[name] for name in ACCUM_FIELDS)

    @staticmethod
    def backward(ctx, *grads):  # type: ignore[override]
        valid, tx_id, rx_id, component_id, num_tx, num_rx, coherent, scattering = ctx.gates
        f = ctx.fwd
        g = dict(zip(ACCUM_FIELDS, grads))
        out = deterministic_accumulate_flat_backward(
            valid, tx_id, rx_id, component_id, f["component_field_real"], f["component_field_imag"],
            f["field_total_real"], f["field_total_imag"], f["power_total"], grad_power_total=g["power_total"],
            grad_field_total_real=g["field_total_real"], grad_field_total_imag=g["field_total_imag"],
            grad_component_power=g["component_power"], grad_component_field_real=g["component_field_real"],
            grad_component_field_imag=g["component_field_imag"], num_tx=num_tx, num_rx=num_rx, coherent=coherent,
            scattering_combine_domain=scattering)
        return (out["grad_path_gain"], out["grad_field_real"], out["grad_field_imag"]) + (None,) * 8

    @staticmethod
    def jvp(ctx, t_gain, t_real, t_imag, *_):  # type: ignore[override]
        valid, tx_id, rx_id, component_id, num_tx, num_rx, coherent, scattering = ctx.gates
        f = ctx.fwd
        out = deterministic_accumulate_flat_jvp(
            valid, tx_id, rx_id, component_id, f["component_field_real"], f["component_field_imag"],
            f["power_total"], tangent_path_gain=t_gain, tangent_field_real=t_real, tangent_field_imag=t_imag,
            num_tx=num_tx, num_rx=num_rx, coherent=coherent, scattering_combine_domain=scattering)
        return tuple(out[name] for name in ACCUM_FIELDS)


def deterministic_accumulate_flat_ad(valid, tx_id, rx_id, component_id, path_gain, field_real, field_imag, *,
                                     num_tx: int, num_rx: int, coherent: bool,
                                     scattering_combine_domain: int = 0) -> dict[str, torch.Tensor]:
    """Differentiable accumulation (reference kernels/deterministic.py:370-390):
    float32 or float64 rows, reverse- and forward-mode."""
    outs = _DeterministicAccumulateFlatAdFunction.apply(path_gain, field_real, field_imag, valid, tx_id, rx_id,
                                                        component_id, int(num_tx), int(num_rx), bool(coherent),
                                                        int(scattering_combine_domain))
    return dict(zip(ACCUM_FIELDS, outs))


__all__ = ["deterministic_accumulate_flat", "deterministic_accumulate_flat_ad", "deterministic_accumulate_flat_backward",
           "deterministic_accumulate_flat_fwd64", "deterministic_accumulate_flat_jvp"]
