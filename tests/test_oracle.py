"""Pinning the CPU oracle of the kinematic sum (CPU only).

1. Golden vectors: outputs of the reference's own compute_scattering_cffi
   (Beam.py:663-816, compiled from its source text by oracle/build_ref.py)
   on seeded inputs, every flag combination, two polarization states.
2. Closed forms independent of the oracle's code: the forward-direction
   single-atom amplitude r_e (f0(0)+f'+if'') e^{ikR}; the Laue interference
   function of a perfect cubic crystal in the far field; the Au table
   values quoted from the data files.
3. The long-double C restatement and the long-double NumPy restatement
   agree with each other, and the float64 paths sit within the float64
   phase-rounding floor of them.
"""
import os

import numpy as np
import pytest

from mosaic_amd import tables
from mosaic_amd.detector import detector
from oracle import kinematic as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "kin_golden.npz"))
CASES = ("three_atoms", "random64", "forward25", "near30nm")


def _args(name):
    g = GOLD
    pc = g[f"{name}/coords"]
    ny, nz = (int(v) for v in g[f"{name}/shape"])
    amp = g[f"{name}/amp"]
    return dict(pos_m=g[f"{name}/pos"], f0p=g[f"{name}/f0p"], f0z=g[f"{name}/f0z"],
                sanr=g[f"{name}/sanr"], sani=g[f"{name}/sani"], amp_r=amp.real, amp_i=amp.imag,
                cx=pc[0], cy=pc[1], cz=pc[2], k_val=float(g["k"]), Ny=ny, Nz=nz)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("fi", range(5))
@pytest.mark.parametrize("pi", range(2))
def test_c_and_numpy_match_reference_golden(name, fi, pi):
    rf, pol, dec = (bool(v) for v in GOLD["flags"][fi])
    pv = GOLD["pol"][pi]
    kw = dict(_args(name), remove_forward=rf, apply_pol=pol, apply_decay=dec,
              pol_deg=pv[0], pol_e=pv[1:])
    want = GOLD[f"{name}/out_f{fi}_p{pi}"].astype(np.complex128)
    scale = np.max(np.abs(want))
    # Any two float64 evaluations of k (x + r) ~ 5e10 rad differ by ~1e-5 rad
    # per pair (ulp of r at 1 m is 1.1e-16 m, fmod by the double 2 pi adds
    # ~4e-6 rad), so the bound is 5e-5 of the largest pixel: a dropped term or
    # a sign error moves the answer by percents.
    for got in (O.kin_sum_c(**kw, precision="f64"), O.kin_sum_numpy(**kw),
                O.kin_sum_c(**kw, precision="f80")):
        assert np.max(np.abs(got - want)) <= 5e-5 * scale


def test_golden_flags_change_the_answer():
    """Each flag is exercised: the golden outputs differ between flag sets."""
    for fi in range(1, 5):
        changed = [not np.allclose(GOLD[f"{n}/out_f0_p0"], GOLD[f"{n}/out_f{fi}_p0"],
                                   rtol=1e-5, atol=0) for n in CASES]
        assert any(changed), fi


def test_long_double_restatements_agree():
    kw = dict(_args("random64"), remove_forward=True, apply_pol=True, apply_decay=True,
              pol_deg=0.3, pol_e=(0.0, 0.0, 1.0))
    a = O.kin_sum_c(**kw, precision="f80")
    b = O.kin_sum_numpy(**kw, dtype=np.longdouble).astype(np.complex128)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
    c = O.kin_sum_c(**kw, precision="f64")
    assert np.linalg.norm(c - a) / np.linalg.norm(a) < 2e-5


def test_forward_single_atom_closed_form():
    """Atom at the origin, pixel on +x: Q = 0, P = 1, r = R0."""
    db = tables.parse_f0_db_all()
    p = db["Au"]
    f00 = float(p[5] + p[:5].sum())
    anom = tables.get_f1f2_from_params(10000.0, tables.parse_f1f2_db_all()["Au"])
    k = 2 * np.pi / tables.wavelength_from_energy_ev(10000.0)
    R = 0.731
    A = O.kin_sum_c(np.zeros((1, 3)), p[None], np.array([f00]), np.array([anom.real]),
                    np.array([anom.imag]), np.ones(1), np.zeros(1), np.array([R]),
                    np.zeros(1), np.zeros(1), k, False, 1, 1, True, 1.0, (0.0, 1.0, 0.0),
                    True, precision="f80")[0]
    two_pi = 2 * np.arccos(np.longdouble(-1))
    ph = float(np.fmod(np.longdouble(k) * np.longdouble(R), two_pi))
    want = 2.81794092e-15 * (f00 + anom) * np.exp(1j * ph)
    assert abs(A - want) <= 1e-6 * abs(want)


def test_laue_function_far_field():
    """N^3 simple-cubic crystal, rotated so (100) diffracts to the detector centre:
    |A| = r_e |S| sqrt(P) prod_i |sin(N x_i) / sin(x_i)|, x_i = q . a_i / 2,
    q = k (u_hat - x_hat), over a detector spanning the whole Laue peak."""
    from mosaic_amd.sample import orient_for_reflection
    N, a = 5, 3.0e-10
    lam = tables.wavelength_from_energy_ev(10000.0)
    k = 2 * np.pi / lam
    th = np.arcsin(lam / (2 * a))
    d = detector()
    d.create_detector((15, 15), (1.0e8, 1.0e8))           # 1 cm pixels
    d.position_detector_absolute(1.0e10, np.degrees(2 * th), 0.0)
    pc = d.pixel_coordinates / 1e10
    R = orient_for_reflection((1, 0, 0), th)
    g = (np.arange(N) - (N - 1) / 2) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) @ R.T
    n = pos.shape[0]
    p = tables.parse_f0_db_all()["Cu"]
    A = O.kin_sum_c(pos, np.repeat(p[None], n, 0), np.zeros(n), np.zeros(n), np.zeros(n),
                    np.ones(n), np.zeros(n), pc[0], pc[1], pc[2], k, False, 15, 15, True, 1.0,
                    (0.0, 1.0, 0.0), False, precision="f80")
    u = pc / np.linalg.norm(pc, axis=0)
    q = k * (u - np.array([[1.0], [0.0], [0.0]]))
    x = (R.T @ q) * a / 2                                  # q . (R e_i) a / 2
    sx = np.sin(x)
    safe = np.abs(sx) > 1e-12
    D = np.prod(np.where(safe, np.abs(np.sin(N * x) / np.where(safe, sx, 1.0)), N), axis=0)
    Q = np.linalg.norm(q, axis=0)
    s2 = (Q * 1e-10 / (4 * np.pi)) ** 2
    f0 = p[5] + sum(p[i] * np.exp(-p[6 + i] * s2) for i in range(5))
    P = 1.0 - u[1] ** 2
    want = 2.81794092e-15 * f0 * np.sqrt(P) * D
    assert np.max(np.abs(np.abs(A) - want)) <= 1e-6 * want.max()
    c = 7 * 15 + 7                                         # detector centre: the Bragg peak
    assert want[c] > 0.99 * 2.81794092e-15 * f0[c] * np.sqrt(P[c]) * N ** 3
    assert want.min() < 0.2 * want[c]                      # and the fringes are resolved


def test_au_table_values():
    p = tables.parse_f0_db_all()["Au"]
    assert np.allclose(p[:6], [16.777390, 19.317156, 32.979683, 5.595453, 10.576854, -6.279078])
    f00 = tables.build_f0_zero_dict({"Au": p})["Au"]
    assert abs(f00 - 78.967458) < 1e-4
    # Cromer-Liberman Au rows bracketing 10 keV (f1f2_CromerLiberman.dat, Au section)
    tbl = tables.parse_f1f2_db_all()["Au"]
    i = np.searchsorted(tbl[:, 0], 10000.0)
    assert np.allclose(tbl[i - 1], [9959.26, -5.76303, 5.19712], rtol=1e-6)
    assert np.allclose(tbl[i], [10000.7, -5.80842, 5.16251], rtol=1e-6)
    c = tables.get_f1f2_from_params(10000.0, tbl)
    w = (10000.0 - 9959.26) / (10000.7 - 9959.26)
    assert abs(c.real - (-5.76303 + w * (-5.80842 + 5.76303))) < 1e-5
    assert abs(c.imag - (5.19712 + w * (5.16251 - 5.19712))) < 1e-5


def test_reference_build_matches_golden_when_present():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built here")
    rf, pol, dec = (bool(v) for v in GOLD["flags"][2])
    pv = GOLD["pol"][0]
    kw = dict(_args("random64"), remove_forward=rf, apply_pol=pol, apply_decay=dec,
              pol_deg=pv[0], pol_e=pv[1:])
    got = O.ref_kin_sum(**kw, threads=1)
    assert np.array_equal(got.astype(np.complex64), GOLD["random64/out_f2_p0"])
