"""Pinning the propagation oracle and the host rules (CPU only).

* golden vectors of the reference's prop_mul_cpu (Beam.py:4853-4891);
* the pad-size rule and sub-step count worked by hand, and the product's
  host rules equal to the oracle's on a sweep;
* closed forms: a Gaussian beam through a thin lens and free space
  (paraxial q-parameter law), a propagate(z) / propagate(-z) round trip,
  Parseval.
"""
import os

import numpy as np
import pytest

from mosaic_amd import propagation as MP
from oracle import propagation as P

HERE = os.path.dirname(os.path.abspath(__file__))
LAM = 1.2398419843320025e-10
K = 2 * np.pi / LAM


def test_transfer_function_matches_reference_golden():
    g = np.load(os.path.join(HERE, "golden", "prop_golden.npz"))
    n_evan = 0
    for i in range(3):
        H = P.tf_from_k(g[f"c{i}/ky"], g[f"c{i}/kz"], float(g["k"]), float(g[f"c{i}/z"]))
        n_evan += int((np.abs(H) < 0.999).sum())
        got = g[f"c{i}/F"] * H
        want = g[f"c{i}/out"]
        # the reference forms the phase in float32: ~1e-5 of the largest entry
        assert np.max(np.abs(got - want)) <= 2e-5 * np.max(np.abs(want)), i
    assert n_evan > 1000   # the evanescent branch is exercised


def test_pad_rule_by_hand():
    # 100 px at 100 nm, z = 1 cm: sin = lambda / 2d = 6.2e-4, pad = ceil(1.1 z tan / d) = 69
    assert P.choose_optimal_pad(100, 100, 1e-7, 1e-7, LAM, 0.01) == (256, 256)
    assert MP.choose_optimal_pad(100, 100, 1e-7, 1e-7, LAM, 0.01) == (256, 256)
    # 2048 px at 170 nm over 2 m hits the 8192 cap; z_max = 8192 d^2 / lambda = 1.909 m,
    # so two 1 m steps, each padded to 8192
    assert P.choose_optimal_pad(2048, 2048, 170e-9, 170e-9, LAM, 2.0) == (8192, 8192)
    zmax = P.alias_free_step(2048, 2048, 170e-9, 170e-9, LAM, 2.0)
    assert abs(zmax - 8192 * (170e-9) ** 2 / LAM) < 1e-12
    steps = MP.plan_steps(2048, 2048, 170e-9, 170e-9, LAM, 2.0)
    assert [(round(z, 12), a, b) for z, a, b in steps] == [(1.0, 8192, 8192)] * 2
    # zero distance: only the minimum pad factor
    assert P.choose_optimal_pad(300, 200, 1e-7, 1e-7, LAM, 0.0, min_pad_factor=1.5) == (512, 512)


def test_host_rules_match_oracle_sweep():
    rng = np.random.default_rng(3)
    for _ in range(200):
        ny, nz = (int(v) for v in rng.integers(8, 3000, 2))
        dy, dz = 10 ** rng.uniform(-10.5, -5, 2)
        z = float(10 ** rng.uniform(-6, 1) * rng.choice([-1, 1]))
        pf = float(rng.choice([1.0, 1.3, 2.0]))
        assert MP.choose_optimal_pad(ny, nz, dy, dz, LAM, z, min_pad_factor=pf) == \
            P.choose_optimal_pad(ny, nz, dy, dz, LAM, z, min_pad_factor=pf)
        a = MP.alias_free_step(ny, nz, dy, dz, LAM, z, pf)
        b = P.alias_free_step(ny, nz, dy, dz, LAM, z, pf)
        assert a == b


def gaussian_lens_closed_form(w0, f, z, x, y):
    """exp(-a r^2) / (1 + 2 i a z / k), a = 1/w0^2 + i k / (2 f) (lens), after
    propagating z with exp(-i z kt^2 / 2k)."""
    a = 1.0 / w0 ** 2 + (1j * K / (2 * f) if f else 0.0)
    q = 1.0 + 2j * a * z / K
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    return np.exp(-a * r2 / q) / q


def _grid(n, d):
    return (np.arange(n) - (n - 1) / 2.0) * d


def test_gaussian_lens_propagation_closed_form():
    n, d, w0 = 256, 0.2e-6, 4e-6
    x = _grid(n, d)
    E0 = gaussian_lens_closed_form(w0, 0.0, 0.0, x, x)
    lens = {"kind": "lens box", "number": 1, "focal_length": 200.0, "thickness": 1.0}
    E = P.apply_stack([lens, {"kind": "free space", "length": 30.0}], E0, d, d, LAM,
                      padding_mode="constant")
    want = gaussian_lens_closed_form(w0, 0.2, 0.03, x, x)
    assert np.linalg.norm(E - want) / np.linalg.norm(want) < 1e-6


def test_round_trip_and_parseval():
    n, d, w0 = 256, 0.2e-6, 3e-6
    x = _grid(n, d)
    rng = np.random.default_rng(1)
    E0 = gaussian_lens_closed_form(w0, 0.0, 0.0, x, x) * np.exp(
        1j * 0.3 * rng.standard_normal() * x[None, :] / d / n)
    E1 = P.propagate(E0, d, d, LAM, 0.02, padding_mode="constant")
    assert abs(np.sum(np.abs(E1) ** 2) / np.sum(np.abs(E0) ** 2) - 1) < 1e-9
    E2 = P.propagate(E1, d, d, LAM, -0.02, padding_mode="constant")
    assert np.linalg.norm(E2 - E0) / np.linalg.norm(E0) < 1e-9
