"""Topology facades (reference witwin/channel/kernels/topology.py):
``path_los_export`` (:84), the line-of-sight rows of every (transmitter,
receiver) pair on ``wc_path_los_export`` (csrc/wc_fields.hip). Same argument
meaning, output keys and shapes as the reference; no visibility test."""

from __future__ import annotations

import torch

from witwin.channel import _native
from witwin.channel.kernels.fields import _check, _check_positive_frequency, _Keep


def path_los_export(tx_positions, tx_power, rx_positions, tx_polarizations, *,
                    frequency_hz: float) -> dict[str, torch.Tensor]:
    """Rows of every pair, pair index rx * T + tx: ``tx_id`` / ``rx_id`` i32,
    ``path_length_m`` (clamped at 1e-6 m), ``delay_s``, ``path_gain`` =
    P (lambda / 4 pi d)^2 (|p|^2 - (p . k)^2) and ``path_gain_matrix`` [T, R]."""
    _check("tx_positions", tx_positions, torch.float32, 2, (3,))
    _check("tx_power", tx_power, torch.float32, 1)
    _check("rx_positions", rx_positions, torch.float32, 2, (3,))
    _check("tx_polarizations", tx_polarizations, torch.float32, 2, (3,))
    if tx_power.shape[0] != tx_positions.shape[0]:
        raise ValueError("tx_power must have one value per transmitter")
    if tx_polarizations.shape != tx_positions.shape:
        raise ValueError("tx_polarizations must match tx_positions shape")
    _check_positive_frequency(frequency_hz)
    t, r = int(tx_positions.shape[0]), int(rx_positions.shape[0])
    device = tx_positions.device
    n = t * r
    out = {"tx_id": torch.empty(n, dtype=torch.int32, device=device),
           "rx_id": torch.empty(n, dtype=torch.int32, device=device),
           "path_length_m": torch.empty(n, dtype=torch.float32, device=device),
           "delay_s": torch.empty(n, dtype=torch.float32, device=device),
           "path_gain": torch.empty(n, dtype=torch.float32, device=device),
           "path_gain_matrix": torch.empty((t, r), dtype=torch.float32, device=device)}
    if n == 0:
        return out
    p, c = _native.ptr, _Keep()
    _native.call("wc_path_los_export", t, r, c(tx_positions), c(tx_power), c(rx_positions), c(tx_polarizations),
                 float(frequency_hz), p(out["tx_id"]), p(out["rx_id"]), p(out["path_length_m"]), p(out["delay_s"]),
                 p(out["path_gain"]), p(out["path_gain_matrix"]), _native.stream_handle(device))
    return out


__all__ = ["path_los_export"]
