"""GPU tests of the explicit path solve (``witwin.channel.path``): the CFR
kernel against a float64 evaluation of the same sum, and the packed
``PathResult`` of the two-ray link against its closed form and against the
deterministic solver's exported rows (same export, same canonical order)."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest
import torch

from witwin.channel import deterministic, path, workloads

pytestmark = pytest.mark.gpu

C0 = 299_792_458.0


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _cfr_float64(a, tau, valid, freqs, normalize):
    """sum_p a exp(-j 2 pi f (tau - tau0)) in float64 with tau - tau0 rounded
    to float32, as the reference's float32 delay normalization does."""
    a = a.astype(np.complex128)
    tau32 = tau.astype(np.float32)
    out = np.zeros((a.shape[0], a.shape[2], freqs.size), np.complex128)
    for b in range(a.shape[0]):
        v = valid[b]
        t0 = np.float32(tau32[b][v].min()) if (normalize and v.any()) else np.float32(0.0)
        for p in np.nonzero(v)[0]:
            t = np.float64(np.float32(tau32[b, p] - t0))
            out[b] += a[b, p][:, None] * np.exp(-2j * np.pi * freqs[None, :] * t)
    return out


@pytest.mark.parametrize("mode", ["uniform", "list", "raw_delays"])
def test_cfr_kernel_matches_float64(device, mode):
    _say(f"case: CFR kernel, mode {mode}")
    rng = np.random.default_rng(11)
    b, p, s = 37, 9, 2
    a = (rng.normal(size=(b, p, s)) + 1j * rng.normal(size=(b, p, s))).astype(np.complex64)
    tau = rng.uniform(1e-7, 3e-6, (b, p)).astype(np.float32)
    valid = rng.uniform(size=(b, p)) < 0.7
    valid[3] = False                     # a pair without paths
    f_count = 1500                       # not a multiple of the 1024-subcarrier block
    if mode == "list":
        freqs = np.sort(rng.uniform(27.9e9, 28.1e9, f_count))
        f_t = torch.as_tensor(freqs).to(device)
    else:
        df = 100e6 / 1024
        freqs = 28e9 + df * (np.arange(f_count) - f_count // 2)
        f_t = torch.as_tensor(freqs)       # host tensor: detected as a uniform grid
    normalize = mode != "raw_delays"
    got = path.paths_cfr(torch.as_tensor(a).to(device), torch.as_tensor(tau).to(device),
                         torch.as_tensor(valid).to(device), f_t, normalize_delays=normalize).cpu().numpy()
    ref = _cfr_float64(a, tau, valid, freqs, normalize)
    scale = np.abs(a).sum(axis=1)[:, :, None] + 1e-30      # sum_p |a| per (pair, time)
    err = np.abs(got - ref) / scale
    _say(f"max |H - H64| / sum|a|: {err.max():.2e}")
    assert err.max() <= 2e-6
    assert np.all(got[3] == 0)


def test_cfr_kernel_empty_inputs(device):
    out = path.paths_cfr(torch.zeros((4, 0, 1), dtype=torch.complex64, device=device),
                         torch.zeros((4, 0), device=device), torch.zeros((4, 0), dtype=torch.bool, device=device),
                         torch.linspace(1e9, 2e9, 16))
    torch.cuda.synchronize()
    assert out.shape == (4, 1, 16) and bool((out == 0).all())


def test_path_solve_two_ray_closed_form(device):
    _say("case: path.solve two-ray 4x4 grid, 3.5 GHz, los + ground reflection")
    scene = workloads.two_ray_scene(shape=(4, 4), spacing=3.0)
    res = path.solve(scene, path.Config(max_depth=1, components={"los", "reflection"}),
                     reference_frequency_hz=3.5e9)
    assert res.a.shape == (16, 1, 1, 1, 2, 1)
    assert bool(res.valid.all())
    assert res.num_paths.reshape(-1).tolist() == [2] * 16
    tx = np.array([0.0, 0.0, 10.0])
    image = np.array([0.0, 0.0, -10.0])
    origin = np.array([-4.5, -4.5, 1.5])
    rx = np.array([origin + [3.0 * r, 3.0 * c, 0.0] for r in range(4) for c in range(4)])
    d_los = np.linalg.norm(rx - tx, axis=1)
    d_ref = np.linalg.norm(rx - image, axis=1)
    tau = res.tau.reshape(16, 2).cpu().numpy().astype(np.float64)
    # canonical order inside a pair: depth 0 (LoS) then depth 1 (reflection)
    assert np.allclose(tau[:, 0], d_los / C0, rtol=1e-6)
    assert np.allclose(tau[:, 1], d_ref / C0, rtol=1e-6)
    itype = res.interaction_type.reshape(16, 2, 1).cpu().numpy()
    assert (itype[:, 0, 0] == 0).all() and (itype[:, 1, 0] == int(path.InteractionType.REFLECTION)).all()
    # the specular point lies on the ground where the image ray crosses z = 0
    hit = res.position.reshape(16, 2, 1, 3).cpu().numpy()[:, 1, 0]
    s = 10.0 / (10.0 + 1.5)
    expect = tx + s * (rx - tx) * np.array([1, 1, 0]) + np.array([0, 0, -10.0])
    assert np.allclose(hit[:, :2], expect[:, :2], atol=1e-4) and np.allclose(hit[:, 2], 0.0, atol=1e-4)
    theta_t = res.theta_t.reshape(16, 2).cpu().numpy()
    dep = rx - tx
    assert np.allclose(theta_t[:, 0], np.arccos(dep[:, 2] / np.linalg.norm(dep, axis=1)), atol=1e-5)
    theta_r = res.theta_r.reshape(16, 2).cpu().numpy()
    arr = -(rx - tx)
    assert np.allclose(theta_r[:, 0], np.arccos(arr[:, 2] / np.linalg.norm(arr, axis=1)), atol=1e-5)
    # the same rows as the deterministic export, coefficient for coefficient
    det = deterministic.solve(scene, deterministic.Config(max_depth=1, components={"los", "reflection"},
                                                          export_paths=True), reference_frequency_hz=3.5e9)
    coeff = det.paths.coefficient.cpu().numpy()
    assert np.array_equal(res.a.reshape(-1).cpu().numpy(), coeff)
    # CFR of the packed result equals the float64 sum over its own paths
    freqs = 3.5e9 + (100e6 / 64) * (np.arange(64) - 32)
    h = res.cfr(torch.as_tensor(freqs)).reshape(16, 64).cpu().numpy()
    a = res.a.reshape(16, 2, 1).cpu().numpy()
    ref = _cfr_float64(a, res.tau.reshape(16, 2).cpu().numpy(), np.ones((16, 2), bool), freqs, True)[:, 0]
    assert np.abs(h - ref).max() <= 2e-6 * np.abs(a).sum(axis=1).max()


def test_path_solve_canyon_matches_export_and_max_paths(device):
    _say("case: path.solve urban canyon 6x6 grid at 20 m, depth 2, los+refl+diff, max_paths 4")
    scene = workloads.urban_canyon_scene(seed=2, shape=(6, 6), spacing=20.0)
    cfg = dict(max_depth=2, components={"los", "reflection", "diffraction"})
    res = path.solve(scene, path.Config(**cfg), reference_frequency_hz=28e9)
    det = deterministic.solve(scene, deterministic.Config(export_paths=True, **cfg), reference_frequency_hz=28e9)
    rows = det.paths
    counts = torch.bincount(rows.rx_id.long(), minlength=36).cpu()
    assert res.num_paths.reshape(-1).cpu().tolist() == counts.tolist()
    assert torch.equal(res.tau[res.valid].cpu(), rows.delay_s.cpu())
    assert torch.equal(res.a[res.valid].reshape(-1).cpu(), rows.coefficient.cpu())
    # diffraction rows carry the edge id in the first object slot
    itype = res.interaction_type[res.valid].cpu()
    objects = res.primitive_id[res.valid].cpu()
    diff = rows.component_id.cpu() == 2
    assert (itype[diff, 0] == int(path.InteractionType.DIFFRACTION)).all()
    assert torch.equal(objects[diff, 0], rows.edge_id.cpu()[diff])
    capped = path.solve(scene, path.Config(max_paths=4, **cfg), reference_frequency_hz=28e9)
    assert capped.max_num_paths == min(4, res.max_num_paths)
    assert torch.equal(capped.num_paths.cpu(), torch.clamp(res.num_paths.cpu(), max=4))
    keep = capped.valid.cpu()
    assert torch.equal(capped.tau.cpu()[keep], res.tau.cpu()[:, :, :, :, :capped.max_num_paths][keep])
