"""GPU tests of the row-level field and layer-stack operators
(witwin.channel.kernels.fields / .materials on csrc/wc_rows.hip and
csrc/wc_fields.hip).

* Layer stacks: the reference's own analytic cases
  (tests/kernels/test_em_layer_stack.py:110-357) against the oracle's
  characteristic-matrix stack, with the reference's tolerances.
* Transmission rows: the reference's cases
  (tests/kernels/test_field_transmission_sequence.py:67-172) and random
  multi-wall rows against the float64 oracle rows.
* Adjoint and tangent companions of free space, reflection sequences,
  transmission sequences and layer stacks: along random directions of each
  input kind, <J^T g, v> from the backward op and <g, J v> from the jvp op
  both equal the central difference of the float64 oracle loss <g, f(x)>;
  the differentiable ``*_ad`` forms carry the same derivatives under torch
  autograd (reverse and forward mode).
"""

from __future__ import annotations

import cmath
import math
import sys

import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

import oracle
from witwin.channel.kernels import fields as F
from witwin.channel.kernels import materials as M

pytestmark = pytest.mark.gpu
C0 = 299792458.0


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _csr(materials, device):
    offsets, counts, rows = [], [], []
    for layers in materials:
        offsets.append(len(rows))
        counts.append(len(layers))
        rows.extend(layers)
    arr = np.array(rows, np.float32).reshape(-1, 4)
    t = lambda a, d: torch.as_tensor(np.asarray(a), dtype=d).to(device)  # noqa: E731
    return {"layer_offset": t(offsets, torch.int32), "layer_count": t(counts, torch.int32),
            "layer_thickness_m": t(arr[:, 0], torch.float32), "layer_eps_r": t(arr[:, 1], torch.float32),
            "layer_sigma_e": t(arr[:, 2], torch.float32), "layer_mu_r": t(arr[:, 3], torch.float32)}


def _f32(materials):
    """Layer values as the GPU sees them (float32-rounded)."""
    return [[tuple(float(np.float32(x)) for x in layer) for layer in layers] for layers in materials]


def _eval(device, cos, mids, materials, f):
    return M.em_layer_stack_eval(torch.tensor(cos, dtype=torch.float32, device=device),
                                 torch.tensor(mids, dtype=torch.int32, device=device), **_csr(materials, device),
                                 frequency_hz=f)


def _cx(out, name, i=0):
    return complex(out[f"{name}_real"][i].item(), out[f"{name}_imag"][i].item())


# ----------------------------------------------------------------------------- layer stacks
def test_layer_stack_matches_oracle_on_reference_cases(device):
    _say("case: layer stack, reference analytic cases vs characteristic-matrix oracle")
    # normal incidence, lossless (reference :110-137)
    f, layer = 3.0e9, (0.03, 4.0, 0.0, 1.0)
    out = _eval(device, [1.0], [0], [[layer]], f)
    n, k0 = 2.0, 2 * math.pi * f / C0
    r12 = (1 - n) / (1 + n)
    p = cmath.exp(-1j * n * k0 * layer[0])
    r_expected = r12 * (1 - p * p) / (1 - r12 * r12 * p * p)
    for pol, name in ((0, "te"), (1, "tm")):
        r_o, t_o = oracle.layer_stack([layer], 1.0, f, pol)
        assert abs(_cx(out, f"r_{name}") - r_o) <= 2e-5 * max(abs(r_o), 1.0)
        assert abs(_cx(out, f"t_{name}") - t_o) <= 2e-5 * max(abs(t_o), 1.0)
        assert abs(_cx(out, f"r_{name}") - r_expected) <= 5e-5
    assert abs(out["cap_R_te"][0].item() + out["cap_T_te"][0].item() - 1.0) <= 1e-5
    # oblique lossy single layer (reference :272-297)
    f, layer, cts = 3.5e9, (0.12, 4.0, 0.025, 1.0), [1.0, 0.85, 0.55, 0.25]
    out = _eval(device, cts, [0] * 4, [[layer]], f)
    for i, ct in enumerate(cts):
        for pol, name in ((0, "te"), (1, "tm")):
            r_o, t_o = oracle.layer_stack([layer], ct, f, pol)
            assert abs(_cx(out, f"r_{name}", i) - r_o) <= 2e-5 * max(abs(r_o), 1.0)
            assert abs(_cx(out, f"t_{name}", i) - t_o) <= 2e-5 * max(abs(t_o), 1.0)


def test_layer_stack_analytic_limits(device):
    _say("case: layer stack Brewster / PEC / vacuum phase / energy / halves / extreme / zero-layer")
    n = 2.0
    cb = 1 / math.sqrt(1 + n * n)
    out = _eval(device, [cb], [0], [[(0.02, n * n, 0.0, 1.0)]], 6e9)
    assert out["cap_R_tm"][0].item() < 1e-8 and out["cap_R_te"][0].item() > 0.05
    assert abs(out["cap_R_tm"][0].item() + out["cap_T_tm"][0].item() - 1) <= 1e-5
    out = _eval(device, [1.0, 0.5], [0, 0], [[(0.1, 1.0, 1e9, 1.0)]], 3e9)
    for name in ("cap_R_te", "cap_R_tm", "cap_T_te", "cap_T_tm"):
        assert torch.isfinite(out[name]).all()
    assert (out["cap_R_te"] > 0.999).all() and (out["cap_R_tm"] > 0.999).all()
    assert (out["cap_T_te"] < 1e-20).all() and (out["cap_T_tm"] < 1e-20).all()
    k0 = 2 * math.pi * 3e9 / C0
    for ct in (1.0, math.cos(math.radians(45))):
        out = _eval(device, [ct], [0], [[(0.5, 1.0, 0.0, 1.0)]], 3e9)
        e = cmath.exp(-1j * k0 * ct * 0.5)
        for name in ("te", "tm"):
            assert abs(_cx(out, f"r_{name}")) < 1e-6
            assert abs(_cx(out, f"t_{name}") - e) < 2e-5 * 1.5
    mats = [[(0.04, 4.0, 0.0, 1.0)], [(0.02, 2.5, 0.0, 1.0), (0.03, 6.0, 0.0, 2.0)]]
    cts = [1.0, 0.9, 0.6, 0.3]
    out = _eval(device, cts + cts, [0] * 4 + [1] * 4, mats, 5e9)
    for name in ("te", "tm"):
        budget = out[f"cap_R_{name}"] + out[f"cap_T_{name}"]
        torch.testing.assert_close(budget, torch.ones_like(budget), rtol=0.0, atol=1e-5)
    out = _eval(device, [1.0, 0.7, 0.1], [0, 0, 0], [[(10.0, 3.0, 1e9, 1.0)]], 2e9)
    for name in M.EM_LAYER_STACK_FIELDS:
        assert torch.isfinite(out[name]).all(), name
    out = _eval(device, [0.8, 0.8], [0, 1], [[(0.1, 4.5, 0.02, 1.0)], [(0.05, 4.5, 0.02, 1.0)] * 2], 4e9)
    for name in ("te", "tm"):
        for q in ("r", "t"):
            a, b = _cx(out, f"{q}_{name}", 0), _cx(out, f"{q}_{name}", 1)
            assert abs(a - b) <= 1e-5 * max(abs(a), 1.0)
    out = _eval(device, [0.7], [0], [[]], 3e9)
    assert abs(_cx(out, "r_te")) == 0.0 and _cx(out, "t_te") == 1 + 0j
    assert out["cap_T_te"][0].item() == 1.0 and out["cap_R_tm"][0].item() == 0.0
    # a material id outside the tables gives zeros
    out = _eval(device, [0.7], [5], [[(0.1, 4.0, 0.0, 1.0)]], 3e9)
    assert all(float(out[k].abs().max()) == 0.0 for k in M.EM_LAYER_STACK_FIELDS)


def test_layer_stack_random_rows_match_oracle(device):
    _say("case: layer stack, 400 random rows over 6 random materials vs oracle")
    rng = np.random.default_rng(3)
    mats = _f32([[(rng.uniform(0.001, 0.3), rng.uniform(1, 9), rng.choice([0.0, rng.uniform(0, 0.3)]),
                   rng.choice([1.0, rng.uniform(1, 2)])) for _ in range(rng.integers(1, 5))] for _ in range(6)])
    cts = rng.uniform(0.02, 1.0, 400).astype(np.float32)
    mids = rng.integers(0, 6, 400)
    f = 5.8e9
    out = _eval(device, cts, mids, mats, f)
    host = {k: v.cpu().numpy() for k, v in out.items()}
    worst = 0.0
    for i in range(400):
        for pol, name in ((0, "te"), (1, "tm")):
            r_o, t_o = oracle.layer_stack(mats[mids[i]], float(cts[i]), f, pol)
            r = complex(host[f"r_{name}_real"][i], host[f"r_{name}_imag"][i])
            t = complex(host[f"t_{name}_real"][i], host[f"t_{name}_imag"][i])
            worst = max(worst, abs(r - r_o), abs(t - t_o))
            assert abs(host[f"cap_R_{name}"][i] - abs(r_o) ** 2) <= 5e-5
    _say(f"  worst |delta r|, |delta t| = {worst:.2e}")
    assert worst <= 5e-5


def test_layer_stack_companions_schema_and_facade_errors(device):
    _say("case: layer stack backward/jvp schema and facade rejections")
    cos = torch.tensor([0.7], device=device)
    mid = torch.tensor([0], dtype=torch.int32, device=device)
    layers = _csr([[(0.1, 4.0, 0.02, 1.0)]], device)
    primal = M.em_layer_stack_eval(cos, mid, **layers, frequency_hz=3e9)
    grads = tuple(torch.ones_like(primal[k]) if i % 2 == 0 else None for i, k in enumerate(M.EM_LAYER_STACK_FIELDS))
    back = M.em_layer_stack_backward(cos, mid, **layers, grad_outputs=grads, frequency_hz=3e9, need_cos_theta=True,
                                     need_layers=True, need_frequency=True)
    assert set(back) == {"grad_cos_theta", "grad_layer_thickness_m", "grad_layer_eps_r", "grad_layer_sigma_e",
                         "grad_frequency"}
    assert back["grad_cos_theta"].shape == cos.shape and back["grad_frequency"].shape == (1,)
    assert back["grad_layer_eps_r"].shape == layers["layer_eps_r"].shape
    jvp = M.em_layer_stack_jvp(cos, mid, **layers, frequency_hz=3e9, tangent_cos_theta=torch.full_like(cos, 0.1),
                               tangent_layer_thickness=torch.full_like(layers["layer_thickness_m"], 0.01),
                               tangent_layer_eps_r=torch.full_like(layers["layer_eps_r"], 0.2),
                               tangent_layer_sigma_e=None, tangent_frequency=1e6)
    assert set(jvp) == set(M.EM_LAYER_STACK_FIELDS) and all(v.shape == cos.shape for v in jvp.values())
    with pytest.raises(ValueError, match="one cotangent slot"):
        M.em_layer_stack_backward(cos, mid, **layers, grad_outputs=(None,), frequency_hz=3e9, need_cos_theta=False,
                                  need_layers=False, need_frequency=False)
    with pytest.raises(ValueError, match="must match cos_theta length"):
        M.em_layer_stack_eval(cos, torch.tensor([0, 0], dtype=torch.int32, device=device), **layers,
                              frequency_hz=3e9)
    with pytest.raises(ValueError, match="frequency_hz must be positive"):
        M.em_layer_stack_eval(cos, mid, **layers, frequency_hz=0.0)


# ----------------------------------------------------------------------------- transmission rows
def _transmission(device, src, tgt, normal, hit, pol, materials, f, material_index=0, path_valid=True):
    d = lambda a, t=torch.float32: torch.tensor(a, dtype=t, device=device)  # noqa: E731
    return F.field_transmission_sequence(d([path_valid], torch.bool), d([src]), d([tgt]), d([[hit]]), d([[normal]]),
                                         d([[material_index]], torch.int32), d([[True]], torch.bool), d([1.0]),
                                         d([pol]), d([pol]), **_csr(materials, device), frequency_hz=f)


def _free(device, src, tgt, pol, f):
    d = lambda a: torch.tensor(a, dtype=torch.float32, device=device)  # noqa: E731
    return F.field_free_space(d([src]), d([tgt]), d([1.0]), d([pol]), d([pol]), frequency_hz=f)


def test_transmission_reference_cases(device):
    _say("case: transmission reference cases (invalid row, vacuum wall, lossy wall, length/delay)")
    out = _transmission(device, [math.nan] * 3, [0, 0, -1.0], [math.nan] * 3, [math.nan] * 3, [0, 1.0, 0],
                        [[(0.1, 4.0, 0.05, 1.0)]], 3.5e9, material_index=2 ** 31 - 1, path_valid=False)
    torch.cuda.synchronize()
    for v in out.values():
        assert torch.count_nonzero(v).item() == 0
    for src, tgt in (([0, 0, 2.0], [0, 0, -1.5]), ([-1.5, 0, 1.5], [1.5, 0, -1.5])):
        through = _transmission(device, src, tgt, [0, 0, 1.0], [0, 0, 0.0], [0, 1.0, 0], [[(0.3, 1.0, 0.0, 1.0)]],
                                3e9)
        free = _free(device, src, tgt, [0, 1.0, 0], 3e9)
        ratio = through["coefficient"] / free["coefficient"]
        assert torch.abs(torch.abs(ratio) - 1.0).item() <= 1e-5
        assert torch.abs(torch.angle(ratio)).item() <= 2e-5
        torch.testing.assert_close(through["field_vector"], free["field_vector"], rtol=1e-4, atol=1e-9)
        torch.testing.assert_close(through["direction"], free["direction"])
        torch.testing.assert_close(through["path_length_m"], free["path_length_m"])
    lossy = [[(0.1, 4.0, 0.05, 1.0)]]
    through = _transmission(device, [0, 0, 2.0], [0, 0, -2.0], [0, 0, 1.0], [0, 0, 0.0], [0, 1.0, 0], lossy, 3.5e9)
    free = _free(device, [0, 0, 2.0], [0, 0, -2.0], [0, 1.0, 0], 3.5e9)
    stack = _eval(device, [1.0], [0], lossy, 3.5e9)
    t_mag = math.hypot(stack["t_te_real"][0].item(), stack["t_te_imag"][0].item())
    assert t_mag < 0.999
    observed = (torch.abs(through["coefficient"]) / torch.abs(free["coefficient"])).item()
    assert abs(observed - t_mag) <= 1e-5 * max(t_mag, 1.0)
    assert through["path_gain"][0].item() < free["path_gain"][0].item()
    src, tgt = [1.0, -2.0, 3.0], [-1.0, 2.0, -1.0]
    length = math.dist(src, tgt)
    through = _transmission(device, src, tgt, [0, 0, 1.0], [0, 0, 0.0], [0, 1.0, 0], [[(0.2, 3.0, 0.01, 1.0)]],
                            2.4e9)
    assert through["path_length_m"][0].item() == pytest.approx(length, rel=1e-6)
    assert through["delay_s"][0].item() == pytest.approx(length / C0, rel=1e-6)


def _random_transmission(rng, n, depth, n_mat, device, lossless_layers=True):
    """Random rows; with lossless_layers some layers have sigma = 0, where the
    clamp max(sigma, 0) has a kink (difference quotients straddling it halve)."""
    low = 0.0 if lossless_layers else 0.01
    mats = _f32([[(rng.uniform(0.005, 0.25), rng.uniform(1, 7), rng.choice([0.0, rng.uniform(0, 0.1)])
                   if lossless_layers else rng.uniform(low, 0.1),
                   rng.choice([1.0, rng.uniform(1, 1.5)])) for _ in range(rng.integers(0, 4))] for _ in range(n_mat)])
    src = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    tgt = (src + rng.uniform(5, 20, (n, 1)) * _units(rng, n)).astype(np.float32)
    nrm = _units(rng, n * depth).reshape(n, depth, 3).astype(np.float32)
    mid = rng.integers(0, n_mat, (n, depth)).astype(np.int32)
    iv = rng.uniform(size=(n, depth)) > 0.25
    pv = rng.uniform(size=n) > 0.1
    txp = _units(rng, n).astype(np.float32)
    rxp = _units(rng, n).astype(np.float32)
    power = rng.uniform(0.5, 2, n).astype(np.float32)
    host = dict(path_valid=pv, source=src, target=tgt, interaction_normals=nrm, interaction_material_id=mid,
                interaction_valid=iv, tx_power=power, tx_polarization=txp, rx_polarization=rxp)
    t = lambda a, d: torch.as_tensor(a, dtype=d).to(device)  # noqa: E731
    dev = dict(path_valid=t(pv, torch.bool), source=t(src, torch.float32), target=t(tgt, torch.float32),
               interaction_positions=t(np.zeros((n, depth, 3), np.float32), torch.float32),
               interaction_normals=t(nrm, torch.float32), interaction_material_id=t(mid, torch.int32),
               interaction_valid=t(iv, torch.bool), tx_power=t(power, torch.float32),
               tx_polarization=t(txp, torch.float32), rx_polarization=t(rxp, torch.float32), **_csr(mats, device))
    return mats, host, dev


def _units(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _trans_args(dev):
    keys = ("path_valid", "source", "target", "interaction_positions", "interaction_normals",
            "interaction_material_id", "interaction_valid", "tx_power", "tx_polarization", "rx_polarization",
            "layer_offset", "layer_count", "layer_thickness_m", "layer_eps_r", "layer_sigma_e", "layer_mu_r")
    return [dev[k] for k in keys]


def _oracle_trans(mats, host, f):
    return oracle.transmission_sequence(host["path_valid"], host["source"], host["target"],
                                        host["interaction_normals"], host["interaction_material_id"],
                                        host["interaction_valid"], host["tx_power"], host["tx_polarization"],
                                        host["rx_polarization"], mats, f)


def test_transmission_random_rows_match_oracle(device):
    _say("case: transmission, 300 random rows, depth 3, 5 materials (some empty) vs float64 oracle")
    rng = np.random.default_rng(11)
    f = 3.5e9
    mats, host, dev = _random_transmission(rng, 300, 3, 5, device)
    out = F.field_transmission_sequence(*_trans_args(dev), frequency_hz=f)
    ref = _oracle_trans(mats, host, f)
    got = {k: v.cpu().numpy() for k, v in out.items()}
    scale = np.abs(ref["coefficient"]).max()
    err = np.abs(got["coefficient"] - ref["coefficient"]).max() / scale
    _say(f"  max |delta coefficient| / max |coefficient| = {err:.2e}")
    assert err <= 2e-5
    np.testing.assert_allclose(got["field_vector"], ref["field_vector"], atol=2e-5 * np.abs(ref["field_vector"]).max())
    np.testing.assert_allclose(got["path_gain"], ref["path_gain"], rtol=1e-4, atol=1e-4 * ref["path_gain"].max())
    np.testing.assert_allclose(got["path_length_m"], ref["path_length_m"], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(got["delay_s"], ref["delay_s"], rtol=1e-6, atol=1e-15)
    np.testing.assert_allclose(got["direction"], ref["direction"], atol=1e-6)
    assert not got["path_gain"][~host["path_valid"]].any()


# ----------------------------------------------------------------------------- AD: directional checks
def _weights(rng, out: dict, keys):
    """Random cotangents, each output scaled to contribute O(1)."""
    g = {}
    for k in keys:
        v = out[k].detach()
        s = float(v.abs().max()) or 1.0
        w = torch.as_tensor(rng.uniform(-1, 1, tuple(v.shape) + ((2,) if v.is_complex() else ())) / s,
                            dtype=torch.float32).to(v.device)
        g[k] = torch.view_as_complex(w.contiguous()) if v.is_complex() else w
    return g


def _pair(g: dict, host: dict, keys) -> float:
    """<g, out> with complex outputs as (re, im) pairs, in float64."""
    total = 0.0
    for k in keys:
        gv = g[k].detach().cpu()
        if gv.is_complex():
            gv = torch.view_as_real(gv)
            hv = np.stack([np.real(host[k]), np.imag(host[k])], -1)
        else:
            hv = np.asarray(host[k], np.float64)
        total += float((gv.double().numpy() * hv).sum())
    return total


def _dot(a, b) -> float:
    if a is None or b is None:
        return 0.0
    a = a.detach().double().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a, np.float64)
    b = b.detach().double().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b, np.float64)
    return float((a * b).sum())


def _agree(name, rev, fwd, fd, tol=2e-3):
    scale = max(abs(fd), 1e-9)
    _say(f"  {name:<22} vjp {rev: .6e}  jvp {fwd: .6e}  fd {fd: .6e}")
    assert abs(rev - fd) <= tol * scale, f"{name}: reverse mode differs from the difference quotient"
    assert abs(fwd - fd) <= tol * scale, f"{name}: forward mode differs from the difference quotient"


_GRAD_NAME = {"path_length_m": "path_length", "delay_s": "delay"}


def _grad_kwargs(g, keys):
    return {f"grad_{_GRAD_NAME.get(k, k)}": g[k] for k in keys}


FS_KEYS = F.FIELD_OUTPUT_FIELDS


def _oracle_free_rows(src, tgt, power, txp, rxp, f):
    rows = [oracle.free_space(src[i], tgt[i], power[i], txp[i], rxp[i], f) for i in range(len(src))]
    d = np.asarray(tgt, np.float64) - np.asarray(src, np.float64)
    length = np.linalg.norm(d, axis=1)
    return {"field_vector": np.array([r["field_vector"] for r in rows]),
            "coefficient": np.array([r["coefficient"] for r in rows]),
            "path_field": np.array([r["path_field"] for r in rows]),
            "path_gain": np.array([r["path_gain"] for r in rows]), "path_length_m": length,
            "delay_s": length / C0, "direction": d / length[:, None]}


def test_free_space_adjoint_and_tangent_match_fd(device):
    _say("case: free-space AD, 24 rows, directions in source / target / frequency")
    rng = np.random.default_rng(21)
    n, f = 24, 3.5e9
    src = rng.uniform(-5, 5, (n, 3))
    tgt = src + rng.uniform(3, 30, (n, 1)) * _units(rng, n)
    src, tgt = src.astype(np.float32).astype(np.float64), tgt.astype(np.float32).astype(np.float64)
    txp, rxp = _units(rng, n).astype(np.float32), _units(rng, n).astype(np.float32)
    power = rng.uniform(0.5, 2, n).astype(np.float32)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32)).to(device)  # noqa: E731
    args = (t(src), t(tgt), t(power), t(txp), t(rxp))
    out = F.field_free_space(*args, frequency_hz=f)
    g = _weights(rng, out, FS_KEYS)
    back = F.field_free_space_backward(*args, frequency_hz=f, need_grad_geometry=True, **_grad_kwargs(g, FS_KEYS))
    for kind in ("source", "target", "frequency"):
        v_src = rng.uniform(-1, 1, (n, 3)) if kind == "source" else np.zeros((n, 3))
        v_tgt = rng.uniform(-1, 1, (n, 3)) if kind == "target" else np.zeros((n, 3))
        v_f = float(rng.uniform(-1, 1)) * f * 1e-6 if kind == "frequency" else 0.0
        h = 1e-5 if kind != "frequency" else 1.0
        L = lambda s: _pair(g, _oracle_free_rows(src + s * h * v_src, tgt + s * h * v_tgt, power, txp, rxp,  # noqa: E731
                                                 f + s * h * v_f), FS_KEYS)
        fd = (L(1) - L(-1)) / (2 * h)
        rev = _dot(back["grad_source"], v_src) + _dot(back["grad_target"], v_tgt) + float(back["grad_frequency"]) * v_f
        jv = F.field_free_space_jvp(*args, frequency_hz=f, tangent_frequency=v_f, tangent_source=t(v_src),
                                    tangent_target=t(v_tgt))
        fwd = _pair(g, {k: jv[k].cpu().numpy() for k in FS_KEYS}, FS_KEYS)
        _agree(kind, rev, fwd, fd)


def _oracle_refl_rows(src, tgt, pos, nrm, power, txp, rxp, eps, sig, mu, gain, th, f):
    n = len(src)
    rows = [oracle.reflection_sequence(src[i], tgt[i], pos[i], nrm[i], power[i], txp[i], rxp[i], eps[i], sig[i],
                                       mu[i], gain[i], th[i], f) for i in range(n)]
    last = np.asarray(pos, np.float64)[:, -1]
    d = np.asarray(tgt, np.float64) - last
    length = np.array([r["path_length"] for r in rows])
    return {"field_vector": np.array([r["field_vector"] for r in rows]),
            "coefficient": np.array([r["coefficient"] for r in rows]),
            "path_field": np.array([r["path_field"] for r in rows]),
            "path_gain": np.array([r["path_gain"] for r in rows]), "path_length_m": length, "delay_s": length / C0,
            "direction": d / np.linalg.norm(d, axis=1, keepdims=True)}


def test_reflection_sequence_adjoint_and_tangent_match_fd(device):
    _say("case: reflection-sequence AD, 16 rows, depth 3, every differentiable input kind")
    rng = np.random.default_rng(22)
    n, depth, f = 16, 3, 3.5e9
    src = rng.uniform(-5, 5, (n, 3))
    pos = src[:, None, :] + np.cumsum(rng.uniform(2, 6, (n, depth, 1)) * _units(rng, n * depth).reshape(n, depth, 3),
                                      axis=1)
    tgt = pos[:, -1] + rng.uniform(2, 8, (n, 1)) * _units(rng, n)
    nrm = _units(rng, n * depth).reshape(n, depth, 3)
    q = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    src, pos, tgt, nrm = q(src), q(pos), q(tgt), q(nrm)
    eps = q(rng.uniform(2, 8, (n, depth)))
    sig = q(rng.uniform(0, 0.1, (n, depth)))
    mu = q(np.ones((n, depth)))
    gain = q(rng.uniform(0.5, 1, (n, depth)))
    th = q(rng.uniform(0.01, 0.3, (n, depth)))
    txp, rxp = q(_units(rng, n)), q(_units(rng, n))
    power = q(rng.uniform(0.5, 2, n))
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32)).to(device)  # noqa: E731
    args = [t(x) for x in (src, tgt, pos, nrm, power, txp, rxp, eps, sig, mu, gain, th)]
    out = F.field_reflection_sequence(*args, frequency_hz=f)
    g = _weights(rng, out, FS_KEYS)
    back = F.field_reflection_sequence_backward(*args, frequency_hz=f, need_grad_gain=True, need_grad_geometry=True,
                                                **_grad_kwargs(g, FS_KEYS))
    base = dict(src=src, tgt=tgt, pos=pos, nrm=nrm, eps=eps, sig=sig, gain=gain, th=th)
    grads = dict(src="grad_source", tgt="grad_target", pos="grad_interaction_positions",
                 nrm="grad_interaction_normals", eps="grad_eps_r", sig="grad_sigma_e", gain="grad_gain",
                 th="grad_thickness")
    tans = dict(src="tangent_source", tgt="tangent_target", pos="tangent_interaction_positions",
                nrm="tangent_interaction_normals", eps="tangent_eps_r", sig="tangent_sigma_e", gain="tangent_gain",
                th="tangent_thickness")
    steps = dict(src=1e-5, tgt=1e-5, pos=1e-5, nrm=1e-5, eps=1e-5, sig=1e-6, gain=1e-5, th=1e-6)
    for kind in list(base) + ["frequency"]:
        v = rng.uniform(-1, 1, base[kind].shape) if kind != "frequency" else None
        v_f = float(rng.uniform(-1, 1)) * f * 1e-6 if kind == "frequency" else 0.0
        h = steps.get(kind, 1.0)

        def L(s):
            x = {k: (val + s * h * v if k == kind else val) for k, val in base.items()}
            return _pair(g, _oracle_refl_rows(x["src"], x["tgt"], x["pos"], x["nrm"], power, txp, rxp, x["eps"],
                                              x["sig"], mu, x["gain"], x["th"], f + s * h * v_f), FS_KEYS)

        fd = (L(1) - L(-1)) / (2 * h)
        if kind == "frequency":
            rev = float(back["grad_frequency"]) * v_f
            jv = F.field_reflection_sequence_jvp(*args, frequency_hz=f, tangent_frequency=v_f)
        else:
            rev = _dot(back[grads[kind]], v)
            jv = F.field_reflection_sequence_jvp(*args, frequency_hz=f, **{tans[kind]: t(v)})
        fwd = _pair(g, {k: jv[k].cpu().numpy() for k in FS_KEYS}, FS_KEYS)
        _agree(kind, rev, fwd, fd)


TR_KEYS = F.FIELD_TANGENT_FIELDS


def test_transmission_adjoint_and_tangent_match_fd(device):
    _say("case: transmission AD, 40 rows, depth 3, endpoints / normals / layer tables / frequency")
    rng = np.random.default_rng(23)
    f = 3.5e9
    mats, host, dev = _random_transmission(rng, 40, 3, 4, device, lossless_layers=False)
    args = _trans_args(dev)
    out = F.field_transmission_sequence(*args, frequency_hz=f)
    g = _weights(rng, out, TR_KEYS)
    back = F.field_transmission_sequence_backward(*args, frequency_hz=f, need_grad_geometry=True,
                                                  **_grad_kwargs(g, TR_KEYS))
    lay = np.array([layer for layers in mats for layer in layers], np.float64).reshape(-1, 4)

    def rebuild(col_values):
        """Material lists with layer column values replaced."""
        out_m, i = [], 0
        for layers in mats:
            cur = []
            for layer in layers:
                row = list(layer)
                for col, vals in col_values.items():
                    row[col] = float(vals[i])
                cur.append(tuple(row))
                i += 1
            out_m.append(cur)
        return out_m

    kinds = {"source": ("source", 1e-5), "target": ("target", 1e-5), "normals": ("interaction_normals", 1e-5),
             "thickness": (0, 1e-6), "eps_r": (1, 1e-5), "sigma_e": (2, 1e-6), "frequency": (None, 1.0)}
    for kind, (key, h) in kinds.items():
        v_f = float(rng.uniform(-1, 1)) * f * 1e-6 if kind == "frequency" else 0.0
        if isinstance(key, str):
            v = rng.uniform(-1, 1, host[key].shape)
        elif key is not None:
            v = rng.uniform(-1, 1, lay.shape[0])
        else:
            v = None

        def L(s):
            hx = dict(host)
            mx = mats
            if isinstance(key, str):
                hx[key] = host[key].astype(np.float64) + s * h * v
            elif key is not None:
                mx = rebuild({key: lay[:, key] + s * h * v})
            return _pair(g, _oracle_trans(mx, hx, f + s * h * v_f), TR_KEYS)

        fd = (L(1) - L(-1)) / (2 * h)
        tan = {}
        if kind == "frequency":
            rev = float(back["grad_frequency"]) * v_f
        elif kind in ("source", "target"):
            rev = _dot(back[f"grad_{kind}"], v)
            tan[f"tangent_{kind}"] = torch.as_tensor(v, dtype=torch.float32).to(device)
        elif kind == "normals":
            rev = _dot(back["grad_interaction_normals"], v)
            tan["tangent_interaction_normals"] = torch.as_tensor(v, dtype=torch.float32).to(device)
        else:
            name = {"thickness": "layer_thickness_m", "eps_r": "layer_eps_r", "sigma_e": "layer_sigma_e"}[kind]
            rev = _dot(back[f"grad_{name}"], v)
            tan[f"tangent_{name}"] = torch.as_tensor(v, dtype=torch.float32).to(device)
        jv = F.field_transmission_sequence_jvp(*args, frequency_hz=f, tangent_frequency=v_f, **tan)
        fwd = _pair(g, {k: jv[k].cpu().numpy() for k in TR_KEYS}, TR_KEYS)
        _agree(kind, rev, fwd, fd)


def test_layer_stack_adjoint_and_tangent_match_fd(device):
    _say("case: layer-stack AD, 60 rows over 4 materials, cos / thickness / eps / sigma / frequency")
    rng = np.random.default_rng(24)
    f = 4.2e9
    mats = _f32([[(rng.uniform(0.01, 0.2), rng.uniform(1.5, 7), rng.uniform(0, 0.1), 1.0)
                  for _ in range(rng.integers(1, 4))] for _ in range(4)])
    cts = rng.uniform(0.1, 0.99, 60).astype(np.float32).astype(np.float64)
    mids = rng.integers(0, 4, 60)
    cos = torch.as_tensor(cts, dtype=torch.float32).to(device)
    mid = torch.as_tensor(mids, dtype=torch.int32).to(device)
    csr = _csr(mats, device)
    out = M.em_layer_stack_eval(cos, mid, **csr, frequency_hz=f)
    g = {k: torch.as_tensor(rng.uniform(-1, 1, 60), dtype=torch.float32).to(device) for k in M.EM_LAYER_STACK_FIELDS}
    back = M.em_layer_stack_backward(cos, mid, **csr, grad_outputs=tuple(g[k] for k in M.EM_LAYER_STACK_FIELDS),
                                     frequency_hz=f, need_cos_theta=True, need_layers=True, need_frequency=True)
    lay = np.array([layer for layers in mats for layer in layers], np.float64)
    gh = {k: g[k].double().cpu().numpy() for k in g}

    def L(ct, layer_vals, freq):
        total, i0 = 0.0, 0
        mlist = []
        for layers in mats:
            mlist.append([tuple(layer_vals[i0 + j]) for j in range(len(layers))])
            i0 += len(layers)
        for i in range(60):
            for pol, name in ((0, "te"), (1, "tm")):
                r, t = oracle.layer_stack(mlist[mids[i]], ct[i], freq, pol)
                total += (gh[f"r_{name}_real"][i] * r.real + gh[f"r_{name}_imag"][i] * r.imag
                          + gh[f"t_{name}_real"][i] * t.real + gh[f"t_{name}_imag"][i] * t.imag
                          + gh[f"cap_R_{name}"][i] * abs(r) ** 2 + gh[f"cap_T_{name}"][i] * abs(t) ** 2)
        return total

    for kind, col, h in (("cos_theta", None, 1e-6), ("thickness", 0, 1e-6), ("eps_r", 1, 1e-5),
                         ("sigma_e", 2, 1e-6), ("frequency", None, 1.0)):
        v_c = rng.uniform(-1, 1, 60) if kind == "cos_theta" else np.zeros(60)
        v_l = np.zeros_like(lay)
        if col is not None:
            v_l[:, col] = rng.uniform(-1, 1, lay.shape[0])
        v_f = float(rng.uniform(-1, 1)) * f * 1e-6 if kind == "frequency" else 0.0
        fd = (L(cts + h * v_c, lay + h * v_l, f + h * v_f) - L(cts - h * v_c, lay - h * v_l, f - h * v_f)) / (2 * h)
        rev = (_dot(back["grad_cos_theta"], v_c) + _dot(back["grad_layer_thickness_m"], v_l[:, 0])
               + _dot(back["grad_layer_eps_r"], v_l[:, 1]) + _dot(back["grad_layer_sigma_e"], v_l[:, 2])
               + float(back["grad_frequency"]) * v_f)
        tt = lambda a: torch.as_tensor(a, dtype=torch.float32).to(device)  # noqa: E731
        jv = M.em_layer_stack_jvp(cos, mid, **csr, frequency_hz=f, tangent_cos_theta=tt(v_c),
                                  tangent_layer_thickness=tt(v_l[:, 0]), tangent_layer_eps_r=tt(v_l[:, 1]),
                                  tangent_layer_sigma_e=tt(v_l[:, 2]), tangent_frequency=v_f)
        fwd = sum(_dot(jv[k], g[k]) for k in M.EM_LAYER_STACK_FIELDS)
        _agree(kind, rev, fwd, fd)


def test_ad_forms_carry_the_companion_derivatives(device):
    _say("case: *_ad forms under torch autograd, reverse and forward mode")
    rng = np.random.default_rng(25)
    mats, host, dev = _random_transmission(rng, 20, 2, 3, device)
    f = torch.tensor(3.5e9, dtype=torch.float64)
    eps = dev["layer_eps_r"].clone().requires_grad_(True)
    src = dev["source"].clone().requires_grad_(True)
    freq = f.clone().requires_grad_(True)
    args = _trans_args(dev)
    args[1], args[13] = src, eps
    out = F.field_transmission_sequence_ad(*args, frequency=freq)
    w = _weights(rng, out, F.FIELD_OUTPUT_FIELDS)
    loss = sum((torch.view_as_real(out[k]) * torch.view_as_real(w[k])).sum() if out[k].is_complex()
               else (out[k] * w[k]).sum() for k in F.FIELD_OUTPUT_FIELDS)
    loss.backward()
    back = F.field_transmission_sequence_backward(*_trans_args(dev), frequency_hz=3.5e9, need_grad_geometry=True,
                                                  **_grad_kwargs(w, TR_KEYS))
    torch.testing.assert_close(eps.grad, back["grad_layer_eps_r"], rtol=1e-5, atol=1e-7)
    assert float(freq.grad) == pytest.approx(float(back["grad_frequency"]), rel=1e-5)
    # the source gradient also carries the direction cotangent, (I - d d^T) g_dir / L
    d, length = out["direction"].detach(), out["path_length_m"].detach()
    gd = w["direction"]
    m = (gd - d * (d * gd).sum(-1, keepdim=True)) / length.clamp_min(1e-30)[:, None]
    m = torch.where(length[:, None] > 0, m, torch.zeros_like(m))
    torch.testing.assert_close(src.grad, back["grad_source"] - m, rtol=1e-4, atol=1e-6)
    # forward mode through the reflection form equals the jvp companion
    n, depth = 8, 2
    base = [torch.as_tensor(x, dtype=torch.float32).to(device) for x in (
        rng.uniform(-3, 3, (n, 3)), rng.uniform(6, 9, (n, 3)), rng.uniform(2, 4, (n, depth, 3)),
        _units(rng, n * depth).reshape(n, depth, 3), np.ones(n), _units(rng, n), _units(rng, n),
        rng.uniform(2, 6, (n, depth)), rng.uniform(0, 0.05, (n, depth)), np.ones((n, depth)),
        np.ones((n, depth)), rng.uniform(0.05, 0.2, (n, depth)))]
    t_eps = torch.as_tensor(rng.uniform(-1, 1, (n, depth)), dtype=torch.float32).to(device)
    with fwAD.dual_level():
        eps_d = fwAD.make_dual(base[7], t_eps)
        res = F.field_reflection_sequence_ad(*base[:7], eps_d, *base[8:], frequency=3.5e9)
        tangent = fwAD.unpack_dual(res["coefficient"]).tangent.clone()
    ref = F.field_reflection_sequence_jvp(*base, frequency_hz=3.5e9, tangent_eps_r=t_eps)["coefficient"]
    torch.testing.assert_close(tangent, ref)


def test_path_los_export_matches_closed_form(device):
    _say("case: path_los_export, 7 Tx x 33 Rx against the float64 closed form")
    from witwin.channel.kernels.topology import path_los_export

    rng = np.random.default_rng(31)
    tx = rng.uniform(-20, 20, (7, 3))
    rx = rng.uniform(-20, 20, (33, 3))
    rx[0] = tx[0]  # a coincident pair: the distance clamps at 1e-6 m
    pol = _units(rng, 7)
    power = rng.uniform(0.5, 2.0, 7)
    f = 5.2e9
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32)).to(device)  # noqa: E731
    out = path_los_export(t(tx), t(power), t(rx), t(pol), frequency_hz=f)
    tx32, rx32, pol32, p32 = (np.asarray(a, np.float32).astype(np.float64) for a in (tx, rx, pol, power))
    pair = np.arange(7 * 33)
    ti, ri = pair % 7, pair // 7
    off = tx32[ti] - rx32[ri]
    d = np.maximum(np.linalg.norm(off, axis=1), 1e-6)
    k = off / d[:, None]
    pattern = np.maximum((pol32[ti] ** 2).sum(1) - (pol32[ti] * k).sum(1) ** 2, 0.0)
    gain = p32[ti] * (C0 / f / (4 * np.pi * d)) ** 2 * pattern
    assert np.array_equal(out["tx_id"].cpu().numpy(), ti) and np.array_equal(out["rx_id"].cpu().numpy(), ri)
    np.testing.assert_allclose(out["path_length_m"].cpu().numpy(), d, rtol=1e-6)
    np.testing.assert_allclose(out["delay_s"].cpu().numpy(), d / C0, rtol=1e-6)
    # |p|^2 - (p . k)^2 cancels in float32 when p is nearly along k: the bound
    # carries a term of a few ulp of the uncancelled P (lambda / 4 pi d)^2 |p|^2
    whole = p32[ti] * (C0 / f / (4 * np.pi * d)) ** 2 * (pol32[ti] ** 2).sum(1)
    err = np.abs(out["path_gain"].cpu().numpy().astype(np.float64) - gain)
    assert np.all(err <= 2e-5 * gain + 1e-6 * whole), float((err / (gain + whole)).max())
    np.testing.assert_array_equal(out["path_gain_matrix"].cpu().numpy()[ti, ri], out["path_gain"].cpu().numpy())
