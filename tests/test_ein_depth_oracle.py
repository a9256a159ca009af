"""Pins of the depth-dependent entrance field's oracle (oracle/ein.py) by
closed forms, each of which a plausible slip fails:

* the triangular-shaped-cloud deposit conserves each atom's factor (its
  3 x 3 weights sum to 1) wherever the stencil is inside the grid;
* a uniform slab of N atoms per beam column on the grid nodes gives the
  textbook phase phi = -r_e lambda f N / (du dv) (= -r_e lambda f n t) and
  attenuation tau = r_e lambda f'' N / (du dv) (a sign, lambda power or area
  slip moves it by O(1));
* Ein = E0 at the front face (f = 0), E0 A at the back face (f = 1) and
  E0 A^(1/2) half way;
* the blur (a normalised kernel, circular) keeps the grid's sum; arg A lies in
  (-pi, pi].
"""
import numpy as np

from oracle import ein as OE

R_E_A = 2.81794092e-5


def test_tsc_deposit_conserves_each_atom():
    rng = np.random.default_rng(0)
    NyB, NzB, du, dv, uc, vc = 40, 30, 2.0, 3.0, 19.5, 14.5
    n = 500
    pos = np.zeros((n, 3), np.float32)
    pos[:, 1] = (rng.uniform(2, NyB - 3, n) - uc) * du
    pos[:, 2] = (rng.uniform(2, NzB - 3, n) - vc) * dv
    fr = rng.uniform(1, 80, n).astype(np.float32)
    fi = rng.uniform(0, 5, n).astype(np.float32)
    sre, sim = OE.column_sums(pos, fr, fi, NyB, NzB, du, dv, uc, vc)
    assert np.isclose(sre.sum(dtype=np.float64), fr.sum(dtype=np.float64), rtol=1e-5)
    assert np.isclose(sim.sum(dtype=np.float64), fi.sum(dtype=np.float64), rtol=1e-5)


def _slab(n_x, f1, f2, NyB=16, NzB=12, du=1.5, dv=2.0, lam=1.239842):
    uc, vc = (NyB - 1) / 2.0, (NzB - 1) / 2.0
    iy, iz, ix = np.meshgrid(np.arange(NyB), np.arange(NzB), np.arange(n_x), indexing="ij")
    pos = np.stack([ix.ravel() * 2.0, (iy.ravel() - uc) * du, (iz.ravel() - vc) * dv],
                   axis=1).astype(np.float32)
    n = pos.shape[0]
    sre, sim = OE.column_sums(pos, np.full(n, f1, np.float32), np.full(n, f2, np.float32),
                              NyB, NzB, du, dv, uc, vc)
    A = OE.column_map(sre, sim, du, dv, lam)
    return A, du, dv, lam


def test_uniform_slab_phase_and_attenuation():
    n_x, f1, f2 = 40, 78.2, 4.1
    A, du, dv, lam = _slab(n_x, f1, f2)
    inner = A[2:-2, 2:-2]
    phi = np.angle(inner)
    tau = -np.log(np.abs(inner))
    phi_cf = -R_E_A * lam * f1 * n_x / (du * dv)
    tau_cf = R_E_A * lam * f2 * n_x / (du * dv)
    assert np.allclose(phi, phi_cf, rtol=2e-5) and phi_cf < -0.01
    assert np.allclose(tau, tau_cf, rtol=2e-3)


def test_depth_fraction_front_back_middle():
    n_x, f1, f2 = 30, 60.0, 8.0
    A, du, dv, lam = _slab(n_x, f1, f2)
    tau, phi = OE.sampled_maps(A)
    NyB, NzB = A.shape
    uc, vc = (NyB - 1) / 2.0, (NzB - 1) / 2.0
    E0 = (np.ones((NyB, NzB)) * (0.8 - 0.3j)).astype(np.complex64)
    x = np.array([10.0, 35.0, 60.0], np.float32)               # s_min, middle, s_max
    pos = np.stack([x, np.full(3, (7 - uc) * du), np.full(3, (5 - vc) * dv)],
                   axis=1).astype(np.float32)
    ein = OE.ein_depth(pos, tau, phi, E0, du, dv, uc, vc, 10.0, 60.0)
    a = complex(A[7, 5])
    assert ein[0] == np.complex64(0.8 - 0.3j)
    assert np.isclose(ein[2], (0.8 - 0.3j) * a, rtol=1e-6)
    assert np.isclose(ein[1], (0.8 - 0.3j) * np.sqrt(a), rtol=1e-6)


def test_blur_keeps_sum_and_phase_is_wrapped():
    rng = np.random.default_rng(1)
    sre = rng.uniform(0, 3e5, (24, 20)).astype(np.float32)
    sim = rng.uniform(0, 50, (24, 20)).astype(np.float32)
    A0 = OE.column_map(sre, sim, 1.0, 1.0, 1.24)
    A2 = OE.column_map(sre, sim, 1.0, 1.0, 1.24, kernel_radius=2)
    phi_raw = -(2 * np.pi / 1.24) * (R_E_A * 1.24 ** 2 / (2 * np.pi)) * sre.astype(np.float64)
    assert phi_raw.min() < -np.pi                      # the map wraps
    _, phi0 = OE.sampled_maps(A0)
    assert np.all(phi0 > -np.pi - 1e-6) and np.all(phi0 <= np.pi + 1e-6)
    assert np.allclose(np.exp(1j * phi0), np.exp(1j * phi_raw), atol=1e-3)
    # the blur acts on the unwrapped phase and conserves its sum
    blurred = OE._blur(phi_raw.astype(np.float32), 2)
    assert np.isclose(blurred.sum(dtype=np.float64), phi_raw.sum(), rtol=1e-5)
    assert A2.shape == A0.shape and not np.allclose(A2, A0)
