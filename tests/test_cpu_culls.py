"""The conservative culls checked on the CPU against the exact float tests,
on the device sources themselves (compiled as host C++ through a shim).

Transmitter-leg shadow bins: every bin a triangle sets must be blocked by
that triangle for legs sampled across the bin (float watertight test, the
leg test's end bias); the float pre-test that skips triangles whose in-plane
hull misses the (transmitter, edge) fan triangle must never skip a triangle
that sets a bin."""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from witwin.channel import workloads
from witwin.channel.scene.compiler import compile as compile_scene

HERE = Path(__file__).resolve().parent


def _dump(scene, txs: np.ndarray, out: Path, frequency: float) -> None:
    compiled = compile_scene(scene, reference_frequency_hz=frequency)
    tri = np.zeros((compiled.face_count, 12), np.float32)
    for k in range(3):
        tri[:, 4 * k:4 * k + 3] = compiled.vertices[compiled.faces[:, k]]
    tri[:, 3] = compiled.face_group.astype(np.int32).view(np.float32)
    out.mkdir(parents=True, exist_ok=True)
    tri.tofile(out / "tris.bin")
    compiled.edges.view(np.uint8).tofile(out / "edges.bin")
    np.asarray(txs, np.float32).reshape(-1, 3).tofile(out / "txs.bin")


@pytest.fixture(scope="module")
def checker(tmp_path_factory) -> Path:
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("cull") / "shadow_check"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{HERE / 'cpu_geo'}", str(HERE / "cpu_geo" / "shadow_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    return exe


def _run(exe: Path, data: Path) -> dict[str, int]:
    proc = subprocess.run([str(exe), str(data)], capture_output=True, text=True, timeout=600)
    counts = {k: int(v) for k, v in re.findall(r"(\w+) (\d+)", proc.stdout)}
    assert proc.returncode == 0, proc.stdout + proc.stderr
    return counts


def test_shadow_bins_and_pretest_on_urban_canyon(checker, tmp_path):
    scene = workloads.urban_canyon_scene(seed=2, shape=(2, 2))
    tx_batch, _, _ = workloads.batched_transmitters(seed=3, tx_count=2)
    txs = np.concatenate([[[53.231, 52.013, 25.0]], tx_batch])
    _dump(scene, txs, tmp_path / "canyon", 28e9)
    c = _run(checker, tmp_path / "canyon")
    assert c["violations"] == 0 and c["skipped_with_bins"] == 0
    assert c["set_bins"] > 1000 and c["skipped"] > c["tests"] // 2


def test_shadow_bins_and_pretest_on_heightfield(checker, tmp_path):
    scene = workloads.heightfield_scene(seed=4, size=64, boxes=16, shape=(2, 2))
    _dump(scene, scene.endpoints[0].position.numpy().reshape(1, 3), tmp_path / "terrain", 3.5e9)
    c = _run(checker, tmp_path / "terrain")
    assert c["violations"] == 0 and c["skipped_with_bins"] == 0
    assert c["set_bins"] > 0 and c["skipped"] > c["tests"] // 2
