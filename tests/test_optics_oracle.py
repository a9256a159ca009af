"""Pins of the optics-element and PSF oracle (oracle/optics_elements.py) and
of the detector noise, by closed forms:

* magnifier: unit magnification with nearest sampling is the identity; a
  field linear in (x, y) magnified bilinearly equals the same linear function
  at the mapped coordinates (exactly, with dyadic values), scaled by
  sqrt(R / |Mx My|);
* angular filter: a plane wave at an on-grid spatial frequency is an
  eigenfunction: it leaves multiplied by the mask at its angle (1 or 0 for a
  top-hat, 1 / (1 + (delta / hw)^2) for Darwin);
* zone plate: the amplitude plate transmits the even zones of
  n = r^2 / (lambda f1), f1 = D dr / lambda; the ideal plate is the thin-lens
  phase -k r^2 / (2 f1 / m);
* aperture: the pixels kept are those within the half-width;
* PSF: an isolated bright pixel spreads into the kernel (intensity), and the
  total intensity is kept away from the edges;
* detector noise: the moments the distributions fix.
"""
import numpy as np

from mosaic_amd.detector import detector
from oracle import optics_elements as OE


def _ramp(Ny, Nx):
    y, x = np.mgrid[0:Ny, 0:Nx].astype(np.float64)
    return (1.0 + 0.5 * x - 0.25 * y + 1j * (2.0 - 0.125 * x + 0.5 * y)).astype(np.complex64)


def test_magnifier_identity_and_linear_fields():
    F = _ramp(17, 21)
    assert np.array_equal(OE.bragg_magnifier(F, 1.0, 1.0, order=0), F)
    out = OE.bragg_magnifier(F, 2.0, 4.0, reflectivity=0.25, conserve_energy=True)
    Ny, Nx = F.shape
    cx, cy = (Nx - 1) / 2.0, (Ny - 1) / 2.0
    X = cx + (np.arange(Nx) - cx) / 2.0
    Y = cy + (np.arange(Ny) - cy) / 4.0
    want = (1.0 + 0.5 * X[None, :] - 0.25 * Y[:, None]
            + 1j * (2.0 - 0.125 * X[None, :] + 0.5 * Y[:, None])) * (0.5 / np.sqrt(8.0))
    assert np.allclose(out, want, rtol=0, atol=1e-6)


def test_angular_filter_plane_wave_eigenfunction():
    Ny, Nx, dx, dy, lam = 64, 128, 1e-7, 1e-7, 1.24e-10
    kx_idx = 5
    theta = 2 * np.pi * kx_idx / (Nx * dx) / (2 * np.pi / lam)     # 4.8e-6 rad
    x = np.arange(Nx) * dx
    E = np.broadcast_to(np.exp(2j * np.pi * kx_idx * np.arange(Nx) / Nx), (Ny, Nx)).astype(
        np.complex64)
    for hx_mrad, want in ((theta * 1e3 * 1.01, 1.0), (theta * 1e3 * 0.99, 0.0)):
        out = OE.angular_filter(E, dx, dy, lam, dict(half_angle_x_mrad=hx_mrad))
        assert np.allclose(out, want * E, atol=1e-5)
    hw = 3e-6
    out = OE.angular_filter(E, dx, dy, lam, dict(half_angle_x_mrad=hw * 1e3, rolloff="darwin",
                                                transmission=0.64, phase_shift=0.3))
    g = 0.8 * np.exp(0.3j) / (1.0 + (theta / hw) ** 2)
    assert np.allclose(out, g * E, atol=1e-5)
    del x


def test_zone_plate_zones_and_ideal_phase():
    Ny = Nx = 96
    dx = dy = 20e-9
    lam, drN, D = 1.24e-10, 50.0, 1.5
    zp = dict(outermost_zone_width_nm=drN, diameter_um=D)
    E = np.ones((Ny, Nx), np.complex64)
    out = OE.zone_plate(E, dx, dy, lam, zp)
    f1 = D * 1e-6 * drN * 1e-9 / lam
    x = (np.arange(Nx) - (Nx - 1) / 2.0) * dx
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    n = r2 / (lam * f1)
    safe = (np.abs(n - np.round(n)) > 1e-3) & (np.abs(r2 - (D * 0.5e-6) ** 2) > 1e-15)
    want = ((np.floor(n).astype(int) % 2 == 0) & (r2 <= (D * 0.5e-6) ** 2)).astype(float)
    assert np.array_equal(out.real[safe], want[safe]) and np.all(out.imag == 0)
    ideal = OE.zone_plate(E, dx, dy, lam, dict(zp, zone_plate_type="ideal", order=3))
    inside = r2 < (D * 0.5e-6) ** 2 * 0.99
    k = 2 * np.pi / lam
    assert np.allclose(ideal[inside], np.exp(-1j * k * r2[inside] / (2 * f1 / 3)), atol=2e-3)


def test_aperture_counts():
    E = np.ones((50, 40), np.complex64)
    out = OE.aperture(E, 1e-6, 2e-6, 0.0201, "square")      # half-width 10.05 um
    x = (np.arange(40) - 19.5) * 1e-6
    y = (np.arange(50) - 24.5) * 2e-6
    assert np.count_nonzero(out) == np.sum(np.abs(x) <= 10.05e-6) * np.sum(np.abs(y) <= 10.05e-6)
    out = OE.aperture(E, 1e-6, 1e-6, 0.03, "circular")
    xx = (np.arange(40) - 19.5) * 1e-6
    yy = (np.arange(50) - 24.5) * 1e-6
    assert np.count_nonzero(out) == np.sum(xx[None, :] ** 2 + yy[:, None] ** 2 <= 15e-6 ** 2)


def test_psf_impulse_response():
    k = detector._build_psf_kernel(1.5, 2.5, "pseudo-voigt", rotation=30.0)
    assert np.isclose(k.sum(), 1.0) and k.shape[0] % 2 == 1
    E = np.zeros((40, 50), np.complex64)
    E[20, 25] = 3.0 + 4.0j                                  # |E|^2 = 25, phase atan2(4, 3)
    out = OE.psf(E, k)
    h = k.shape[0] // 2
    patch = np.abs(out[20 - h:20 + h + 1, 25 - h:25 + h + 1]) ** 2
    assert np.allclose(patch, 25.0 * k[::-1, ::-1], rtol=1e-5, atol=1e-12)
    assert np.isclose(np.sum(np.abs(out) ** 2), 25.0, rtol=1e-5)
    assert np.isclose(np.angle(out[20, 25]), np.arctan2(4, 3), atol=1e-6)


def test_detector_noise_moments():
    d = detector()
    d.create_detector((200, 300), (1.0, 1.0))
    base = np.full((300, 200), 2.0 + 1.0j, np.complex64)
    d.input_pixel_values(base)
    d.add_noise("gaussian", sigma=0.1, seed=1)
    n = np.asarray(d.pixel_values) - base
    s = 0.1 * np.sqrt(5.0) / np.sqrt(2.0)
    assert abs(n.real.std() / s - 1) < 0.02 and abs(n.imag.std() / s - 1) < 0.02
    d.input_pixel_values(base)
    d.add_noise("uniform", sigma=0.2, seed=2)
    n = np.asarray(d.pixel_values) - base
    assert np.max(np.abs(n.real)) <= 0.2 * np.sqrt(5.0) + 1e-6
    assert abs(n.real.std() / (0.2 * np.sqrt(5.0) / np.sqrt(3.0)) - 1) < 0.02
    d.input_pixel_values(base)
    d.add_noise("poisson", photon_flux=400.0, seed=3)
    counts = np.abs(np.asarray(d.pixel_values)) ** 2 * (400.0 / 5.0)
    assert abs(counts.mean() / 400.0 - 1) < 0.01 and abs(counts.var() / 400.0 - 1) < 0.05
    assert np.allclose(np.angle(np.asarray(d.pixel_values)), np.angle(base), atol=1e-6)
