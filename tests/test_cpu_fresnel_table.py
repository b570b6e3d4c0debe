"""The Chebyshev table of the Fresnel auxiliary functions behind the device
UTD transition function: its extended-precision reference values against an
independent implementation (scipy.special.fresnel) and the asymptotic series,
and the committed header against the generator's output."""

from __future__ import annotations

import importlib.util
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
from scipy.special import fresnel as sp_fresnel

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "witwin-channel_amd" / "csrc"


def _generator():
    spec = importlib.util.spec_from_file_location("gen_fresnel_aux", CSRC / "gen_fresnel_aux.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _aux_from_scipy(x: float) -> tuple[float, float]:
    s, c = sp_fresnel(x)
    ph = np.pi * x * x / 2
    return ((0.5 - s) * np.cos(ph) - (0.5 - c) * np.sin(ph), (0.5 - c) * np.cos(ph) + (0.5 - s) * np.sin(ph))


@pytest.mark.parametrize("x", [0.0, 0.05, 0.3, 0.9, 1.5, 1.99, 2.0, 2.01, 2.7, 3.5])
def test_reference_aux_matches_scipy(x):
    gen = _generator()
    f, g = gen.aux(x)
    fs, gs = _aux_from_scipy(x)
    # scipy's C, S are accurate to ~1e-16 absolute; f and g are O(0.1) here
    assert abs(float(f) - fs) <= 2e-15 and abs(float(g) - gs) <= 2e-14


@pytest.mark.parametrize("x", [8.0, 11.0, 25.0])
def test_reference_aux_matches_asymptotic_series(x):
    # A&S 7.3.27-7.3.28, the form the device uses above x = 8
    gen = _generator()
    f, g = gen.aux(x)
    z2 = (np.pi * x * x) ** 2
    tf = tg = sf = sg = 1.0
    for m in range(1, 10):
        tf *= -(4 * m - 3) * (4 * m - 1) / z2
        tg *= -(4 * m - 1) * (4 * m + 1) / z2
        sf += tf
        sg += tg
    assert abs(sf / (np.pi * x) / float(f) - 1) < 1e-14
    assert abs(sg / (np.pi ** 2 * x ** 3) / float(g) - 1) < 1e-14


def test_committed_table_is_the_generator_output():
    out = subprocess.run([sys.executable, str(CSRC / "gen_fresnel_aux.py")], capture_output=True, text=True,
                         check=True)
    assert out.stdout == (CSRC / "wc_fresnel_aux_table.h").read_text()
    worst = float(out.stderr.strip().split()[-1])
    assert worst < 4e-15
