"""CPU checks that pin the oracle's characteristic-matrix layer stack
(oracle.layer_stack) and straight transmission rows
(oracle.transmission_sequence) before the GPU tests compare against them.

Each check is one the mathematics fixes independently of the formula being
tested: the normal-incidence impedance formula, a vacuum layer's pure
interior phase, energy conservation of lossless stacks, the Brewster null,
splitting a layer in two, a PEC limit, and agreement with the oracle's
second formulation of the same stack (the backward Rouard recursion in
oracle.cpp) — the two share no code, so a dropped term, a sign slip or a
swapped polarization in either shows. The cases follow the reference's own
tests/kernels/test_em_layer_stack.py:110-300 and
tests/kernels/test_field_transmission_sequence.py:67-172."""

from __future__ import annotations

import cmath
import math

import numpy as np
import pytest

import oracle

C0 = 299792458.0
EPS0 = 8.8541878128e-12


def test_normal_incidence_matches_impedance_formula():
    f, layer, n = 3.0e9, (0.03, 4.0, 0.0, 1.0), 2.0
    k0 = 2 * math.pi * f / C0
    r12 = (1 - n) / (1 + n)
    p = cmath.exp(-1j * n * k0 * layer[0])
    expected = r12 * (1 - p * p) / (1 - r12 * r12 * p * p)
    for pol in (0, 1):
        r, t = oracle.layer_stack([layer], 1.0, f, pol)
        assert abs(r - expected) < 1e-12
        assert abs(abs(r) ** 2 + abs(t) ** 2 - 1) < 1e-12


def test_vacuum_layer_is_pure_interior_phase():
    f, d = 3.0e9, 0.5
    k0 = 2 * math.pi * f / C0
    for ct in (1.0, math.cos(math.radians(45))):
        for pol in (0, 1):
            r, t = oracle.layer_stack([(d, 1.0, 0.0, 1.0)], ct, f, pol)
            assert abs(r) < 1e-12
            assert abs(t - cmath.exp(-1j * k0 * ct * d)) < 1e-12


@pytest.mark.parametrize("pol", [0, 1])
def test_lossless_stacks_conserve_energy(pol):
    for layers in ([(0.04, 4.0, 0.0, 1.0)], [(0.02, 2.5, 0.0, 1.0), (0.03, 6.0, 0.0, 2.0)]):
        for ct in (1.0, 0.9, 0.6, 0.3):
            r, t = oracle.layer_stack(layers, ct, 5e9, pol)
            assert abs(abs(r) ** 2 + abs(t) ** 2 - 1) < 1e-12


def test_brewster_null_is_tm_only():
    n = 2.0
    cb = 1 / math.sqrt(1 + n * n)
    r_tm, _ = oracle.layer_stack([(0.02, n * n, 0.0, 1.0)], cb, 6e9, 1)
    r_te, _ = oracle.layer_stack([(0.02, n * n, 0.0, 1.0)], cb, 6e9, 0)
    assert abs(r_tm) ** 2 < 1e-20
    assert abs(r_te) ** 2 > 0.05


def test_two_half_layers_equal_one_layer():
    for pol in (0, 1):
        a = oracle.layer_stack([(0.1, 4.5, 0.02, 1.0)], 0.8, 4e9, pol)
        b = oracle.layer_stack([(0.05, 4.5, 0.02, 1.0)] * 2, 0.8, 4e9, pol)
        assert abs(a[0] - b[0]) < 1e-12 and abs(a[1] - b[1]) < 1e-12


def test_pec_limit_reflects_everything():
    for ct in (1.0, 0.5):
        for pol in (0, 1):
            r, t = oracle.layer_stack([(0.1, 1.0, 1e9, 1.0)], ct, 3e9, pol)
            assert abs(r) ** 2 > 0.999 and abs(t) ** 2 < 1e-20


def test_single_lossy_layer_te_matches_slab_formula():
    """TE r of one layer equals the thin-slab closed form
    r_int (1 - p^2) / (1 - r_int^2 p^2) (reference
    tests/kernels/test_em_layer_stack.py:285-297) and the oracle's slab_fresnel."""
    f, layer = 3.5e9, (0.12, 4.0, 0.025, 1.0)
    omega = 2 * math.pi * f
    eta = layer[1] - 1j * layer[2] / (omega * EPS0)
    root = oracle._passive_sqrt(eta)
    r_int = (1 - root) / (1 + root)
    p2 = cmath.exp(-2j * (2 * math.pi * layer[0] / (C0 / f) * root))
    legacy = r_int * (1 - p2) / (1 - r_int * r_int * p2)
    r, _ = oracle.layer_stack([layer], 1.0, f, 0)
    assert abs(r - legacy) < 1e-12
    te, _ = oracle.slab_fresnel(1.0, layer[1], layer[2], 1.0, 1.0, layer[0], f)
    assert abs(r - te) < 1e-9


def test_agrees_with_the_rouard_restatement():
    rng = np.random.default_rng(0)
    for _ in range(300):
        arr = np.array([(rng.uniform(0.001, 0.2), rng.uniform(1, 8), rng.choice([0, rng.uniform(0, 0.5)]),
                         rng.choice([1.0, rng.uniform(1, 3)])) for _ in range(rng.integers(1, 4))], np.float32)
        ct, f, pol = float(rng.uniform(0.05, 1)), float(rng.uniform(1e9, 30e9)), int(rng.integers(0, 2))
        _, t = oracle.layer_stack([tuple(map(float, row)) for row in arr], ct, f, pol)
        t2 = oracle.stack_t(ct, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], f, pol)
        assert abs(t - t2) <= 1e-9 * max(abs(t), 1e-3)


def _one_row(src, tgt, normal, layers, f, pol=(0.0, 1.0, 0.0), valid=True, material=0):
    return oracle.transmission_sequence([valid], [src], [tgt], [[normal]], [[material]], [[True]], [1.0], [pol],
                                        [pol], [layers], f)


@pytest.mark.parametrize("src,tgt", [([0, 0, 2.0], [0, 0, -1.5]), ([-1.5, 0, 1.5], [1.5, 0, -1.5])])
def test_transmission_vacuum_wall_equals_free_space(src, tgt):
    out = _one_row(src, tgt, [0, 0, 1.0], [(0.3, 1.0, 0.0, 1.0)], 3e9)
    ref = oracle.free_space(src, tgt, 1.0, [0, 1.0, 0], [0, 1.0, 0], 3e9)
    assert abs(out["coefficient"][0] - ref["coefficient"]) < 1e-12 * abs(ref["coefficient"]) + 1e-15
    assert abs(out["path_length_m"][0] - ref["path_length"]) < 1e-12


def test_transmission_lossy_wall_scales_by_te_magnitude():
    layers = [(0.1, 4.0, 0.05, 1.0)]
    out = _one_row([0, 0, 2.0], [0, 0, -2.0], [0, 0, 1.0], layers, 3.5e9)
    ref = oracle.free_space([0, 0, 2.0], [0, 0, -2.0], 1.0, [0, 1.0, 0], [0, 1.0, 0], 3.5e9)
    _, t_te = oracle.layer_stack(layers, 1.0, 3.5e9, 0)
    assert abs(t_te) < 0.999
    assert abs(abs(out["coefficient"][0]) / abs(ref["coefficient"]) - abs(t_te)) < 1e-12


def test_transmission_invalid_rows_are_zero():
    out = _one_row([math.nan] * 3, [0, 0, -1.0], [math.nan] * 3, [(0.1, 4.0, 0.05, 1.0)], 3.5e9, valid=False,
                   material=2 ** 31 - 1)
    assert all(not np.any(v) for v in out.values())
