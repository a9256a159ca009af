"""GPU optics elements and detector PSF against the oracle's restatements
(oracle/optics_elements.py), element by element and in a full stack.

Bars: the resampling, masks and PSF agree to float32 rounding (relative L2
<= 1e-6); the angular filter, through an FFT round trip on each side, to
5e-6; the stack to 1e-5.
"""
import numpy as np
import pytest
import torch

from mosaic_amd import optics as OP
from mosaic_amd.detector import detector
from oracle import optics_elements as OE

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _field(Ny, Nx, seed=0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:Ny, 0:Nx]
    E = np.exp(-((x - Nx / 2) ** 2 + (y - Ny / 2) ** 2) / (0.2 * Nx * Ny)) * np.exp(
        1j * 0.01 * (x * x - 0.5 * y * y)) + 0.05 * (rng.standard_normal((Ny, Nx))
                                                   + 1j * rng.standard_normal((Ny, Nx)))
    return E.astype(np.complex64)


@pytest.mark.parametrize("kw", [dict(order=0), dict(order=1), dict(order=1, pad_mode="edge"),
                                dict(order=1, conserve_energy=False, reflectivity=0.7,
                                     phase_shift=0.4)])
def test_bragg_magnifier(device, kw):
    F = _field(96, 128)
    elem = dict(magnification_x=1.7, magnification_y=2.3, **kw)
    got = OP.bragg_magnifier(torch.from_numpy(F).to(device), elem).cpu().numpy()
    want = OE.bragg_magnifier(F, 1.7, 2.3, **kw)
    e = rel(got, want)
    print(f"[optics] magnifier {kw}: rel {e:.2e}", flush=True)
    assert e <= 1e-6


@pytest.mark.parametrize("filt", [
    dict(half_angle_x_mrad=0.05),
    dict(half_angle_x_mrad=0.04, rolloff="darwin", transmission=0.8, phase_shift=0.2),
    dict(half_angle_x_mrad=0.06, half_angle_y_mrad=0.02, shape="elliptical", roll_deg=25.0,
         center_x_mrad=0.01),
    dict(half_angle_x_mrad=0.03, mode="1d", rolloff="butterworth", order=3, roll_deg=-10.0),
])
def test_angular_filter(device, filt):
    F = _field(128, 96, seed=1)
    dx, dy, lam = 2e-7, 3e-7, 1.24e-10
    got = OP.angular_filter(torch.from_numpy(F).to(device), dx, dy, filt, lam).cpu().numpy()
    want = OE.angular_filter(F, dx, dy, lam, filt)
    e = rel(got, want)
    print(f"[optics] angular filter {filt}: rel {e:.2e}", flush=True)
    assert e <= 5e-6


@pytest.mark.parametrize("zp", [
    dict(zone_plate_type="amplitude"),
    dict(zone_plate_type="phase", central_stop_fraction=0.2, efficiency=0.3),
    dict(zone_plate_type="ideal", order=3, edge_taper=0.3),
])
def test_zone_plate(device, zp):
    F = _field(160, 128, seed=2)
    z = dict(outermost_zone_width_nm=40.0, diameter_um=2.0, **zp)
    dx, dy, lam = 15e-9, 12e-9, 1.24e-10
    got = OP.zone_plate(torch.from_numpy(F).to(device), dx, dy, z, lam).cpu().numpy()
    want = OE.zone_plate(F, dx, dy, lam, z)
    e = rel(got, want)
    print(f"[optics] zone plate {zp}: rel {e:.2e}", flush=True)
    assert e <= 1e-6


@pytest.mark.parametrize("shape", ["square", "circular"])
def test_aperture(device, shape):
    F = _field(64, 80, seed=3)
    got = OP.aperture(torch.from_numpy(F).to(device), 1e-6, 1.5e-6,
                      dict(width=0.05, type=shape)).cpu().numpy()
    assert np.array_equal(got, OE.aperture(F, 1e-6, 1.5e-6, 0.05, shape))


@pytest.mark.parametrize("kw", [dict(profile="gaussian"),
                                dict(profile="lorentzian", field="amplitude", boundary="reflect"),
                                dict(profile="pseudo-voigt", rotation=35.0, boundary="wrap"),
                                dict(profile="gaussian", boundary="nearest"),
                                dict(profile="gaussian", boundary="mirror")])
def test_detector_psf(device, kw):
    d = detector()
    d.create_detector((48, 40), (7.5e5, 7.5e5))
    E = _field(40, 48, seed=4)
    d.input_pixel_values(E.copy())
    d.apply_psf((1.5, 2.5), **kw)
    got = np.asarray(d.pixel_values).reshape(40, 48)
    k = detector._build_psf_kernel(2.5, 1.5, kw.get("profile", "gaussian"),
                                   kw.get("rotation", 0.0))
    mode = kw.get("boundary", "auto")
    want = OE.psf(E, k, "constant" if mode == "auto" else mode, kw.get("field", "intensity"))
    e = rel(got, want)
    print(f"[optics] psf {kw}: rel {e:.2e}", flush=True)
    assert e <= 1e-6


def test_stack_with_every_element(device):
    """CRL + 20 um free space + aperture + angular filter + zone plate +
    magnifier + free space, through beam.wavefield_propagation's callback,
    against the oracle elements chained with the oracle propagator."""
    from mosaic_amd.beam import beam
    from oracle import propagation as OP_ORACLE
    n, d_ = 128, 50e-9
    lam = 1.2398419843320025e-10
    b = beam()
    b.create_beam(10000.0)
    ops = OP.optics()
    ops.add_CRL_box(1, 200.0, 1.0)
    ops.add_free_space(0.02)
    ops.add_aperture(0.005, "circular")
    ops.add_angular_filter(0.2, rolloff="darwin")
    ops.add_fresnel_zone_plate(30.0, 5.0, zone_plate_type="phase")
    ops.add_bragg_magnifier_2b(1.5, 1.2)
    ops.add_free_space(0.01)
    E = _field(n, n, seed=5)
    prop = b.free_space_propagator()
    got = ops.apply_stack(E, d_, d_, lam, prop)
    # the oracle chain
    want = OP_ORACLE.thin_lens(E, d_, d_, ops.components[0], lam)
    want = OP_ORACLE.propagate(want, d_, d_, lam, 0.02e-3)
    want = OE.aperture(want, d_, d_, 0.005, "circular")
    want = OE.angular_filter(want, d_, d_, lam, ops.components[3])
    want = OE.zone_plate(want, d_, d_, lam, ops.components[4])
    want = OE.bragg_magnifier(want, 1.5, 1.2)
    want = OP_ORACLE.propagate(want, d_, d_, lam, 0.01e-3)
    e = rel(got, want)
    print(f"[optics] full stack: rel {e:.2e}", flush=True)
    assert e <= 1e-5
