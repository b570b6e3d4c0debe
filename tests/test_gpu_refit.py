"""Device refit of a scene with moved vertices (same topology) against a
fresh compile + build of the moved scene: triangle tables, group planes and
edge records, segment visibility and a radiomap."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from witwin.channel import _native, workloads
from witwin.channel.components import component_bits
from witwin.channel.propagation import engine
from witwin.channel.scene.compiler import DeviceScene, compile as compile_scene, with_vertices

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def _moved(vertices: np.ndarray, seed: int, sigma: float) -> np.ndarray:
    """One displacement per distinct position, so welded vertices move together."""
    rng = np.random.default_rng(seed)
    uniq, inv = np.unique(vertices, axis=0, return_inverse=True)
    return (vertices + rng.normal(0.0, sigma, uniq.shape)[inv.reshape(-1)]).astype(np.float32)


@pytest.mark.parametrize("bvh", ["sah", "lbvh"])
@pytest.mark.parametrize("which", ["canyon", "heightfield"])
def test_refit_matches_a_fresh_build(device, which, bvh):
    _say(f"case: device refit, {which}, {bvh} tree, moved vertices vs fresh compile")
    if which == "canyon":
        scene = workloads.urban_canyon_scene(seed=2, shape=(16, 16), spacing=8.0)
        frequency, bits = 28e9, component_bits({"los", "reflection", "diffraction"})
        sigma = 0.05
    else:
        scene = workloads.heightfield_scene(seed=4, size=48, boxes=12, shape=(16, 16))
        frequency, bits = 3.5e9, component_bits({"los", "reflection"})
        sigma = 0.02
    compiled = compile_scene(scene, reference_frequency_hz=frequency)
    moved = _moved(compiled.vertices, seed=5, sigma=sigma)
    fresh_compiled = with_vertices(compiled, moved)
    fresh = DeviceScene(fresh_compiled, device, bvh=bvh)
    dev = DeviceScene(compiled, device, bvh=bvh)
    dev.refit(torch.as_tensor(moved))
    torch.cuda.synchronize()

    # triangle tables: the same float positions
    assert torch.equal(dev.face_tris, fresh.face_tris)
    # planes (float64) and edge records (float32) from the same formulas
    np.testing.assert_allclose(dev.group_plane.cpu().numpy(), fresh.group_plane.cpu().numpy(), rtol=0, atol=1e-12)
    e_dev = dev.edges.cpu().numpy().view(fresh_compiled.edges.dtype)
    e_ref = fresh.edges.cpu().numpy().view(fresh_compiled.edges.dtype)
    for name in e_ref.dtype.names:
        if e_ref[name].dtype.kind == "f":
            np.testing.assert_allclose(e_dev[name], e_ref[name], rtol=2e-6, atol=2e-6, err_msg=name)
        else:
            np.testing.assert_array_equal(e_dev[name], e_ref[name], err_msg=name)

    # segment visibility through the refit tree == the fresh tree (exact triangle tests)
    rng = np.random.default_rng(9)
    lo, hi = moved.min(axis=0), moved.max(axis=0)
    n = 4096
    a = rng.uniform(lo - 5.0, hi + 5.0, (n, 3)).astype(np.float32)
    b = rng.uniform(lo - 5.0, hi + 5.0, (n, 3)).astype(np.float32)
    ta, tb = torch.as_tensor(a).to(device), torch.as_tensor(b).to(device)
    vis = []
    for ds in (dev, fresh):
        out = torch.empty(n, dtype=torch.uint8, device=device)
        _native.call("wc_visibility", ds.ref(), n, _native.ptr(ta), _native.ptr(tb), _native.ptr(out),
                     _native.stream_handle(device))
        vis.append(out.cpu().numpy())
    _say(f"visible segments: {int(vis[1].sum())} of {n}")
    np.testing.assert_array_equal(vis[0], vis[1])

    # radiomap over the refit scene == over the fresh scene
    maps = []
    for comp, ds in ((compiled, dev), (fresh_compiled, fresh)):
        inputs = engine.solve_inputs(comp, ds, components=bits, max_depth=2, coherent=True, frequency_hz=frequency)
        out = engine.radiomap_forward(ds, inputs)
        torch.cuda.synchronize()
        maps.append((out.path_count.cpu().numpy(), out.power_total.cpu().numpy().astype(np.float64)))
    (c0, p0), (c1, p1) = maps
    np.testing.assert_array_equal(c0, c1)
    lit = (p1 > 0) & (p0 > 0)
    err_db = np.abs(10.0 * np.log10(p0[lit] / p1[lit]))
    _say(f"paths {int(c1.sum())}, lit cells {int(lit.sum())}, max dB difference {err_db.max() if err_db.size else 0:.2e}")
    assert (err_db <= 0.01).all()
    assert np.array_equal(p0 > 0, p1 > 0)
