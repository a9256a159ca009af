"""GPU parity of the propagation path against the FP64 oracle and closed forms.

Bar: relative L2 error <= 1e-5 (FP32 FFTs on grids up to 1024^2 sit near
1e-6; the transfer function and lens phases are formed in FP64).
"""
import numpy as np
import pytest
import torch

from mosaic_amd import propagation as MP
from mosaic_amd.beam import beam
from mosaic_amd.detector import detector
from mosaic_amd.optics import optics
from oracle import propagation as P

pytestmark = pytest.mark.gpu
LAM = 1.2398419843320025e-10
TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def smooth_field(ny, nz, seed):
    rng = np.random.default_rng(seed)
    y = np.linspace(-1, 1, ny)
    z = np.linspace(-1, 1, nz)
    env = np.exp(-(y[None, :] ** 2 + z[:, None] ** 2) * 3)
    ph = np.zeros((nz, ny))
    for _ in range(6):
        a, b, c = rng.standard_normal(3)
        ph += a * np.sin(3 * b * y[None, :] + 2 * c * z[:, None])
    return (env * np.exp(1j * ph)).astype(np.complex64)


@pytest.mark.parametrize("method", ["pruned", "fft2d"])
@pytest.mark.parametrize("mode", ["edge", "constant"])
def test_single_step_vs_oracle(device, mode, method):
    E0 = smooth_field(200, 160, 0)
    d, z = 0.1e-6, 0.02
    steps = MP.plan_steps(200, 160, d, d, LAM, z)
    assert len(steps) == 1
    _, Ny2, Nz2 = steps[0]
    got = MP.asp_step(torch.from_numpy(E0).to(device), d, d, LAM, z, Ny2, Nz2,
                      padding_mode=mode, pad_constant=0.25 - 0.5j, method=method).cpu().numpy()
    want = P.asp_step(E0, d, d, LAM, z, Ny2, Nz2, padding_mode=mode, pad_constant=0.25 - 0.5j)
    e = rel(got, want)
    print(f"[parity] asp step {mode} {method} {Ny2}x{Nz2}: rel={e:.3e}")
    assert e <= TOL


# (Ny, Nz, Ny2, Nz2): odd sizes, no padding (R = 1), one axis unpadded,
# lengths at the 64 minimum, Nz a multiple of 128 (the XCD row grouping),
# wide pads (R = 8)
PRUNED_SHAPES = [(200, 160, 512, 256), (64, 64, 64, 64), (100, 3, 256, 64), (37, 256, 64, 256),
                 (130, 384, 1024, 1024), (1000, 128, 8192, 128), (256, 96, 2048, 1024)]


@pytest.mark.parametrize("shape", PRUNED_SHAPES)
@pytest.mark.parametrize("lens", [False, True])
def test_pruned_matches_padded_fft(device, shape, lens):
    """The pruned three-pass step against the padded 2-D hipFFT step, both on
    the GPU, edge and constant padding, with and without the fused lens."""
    Ny, Nz, Ny2, Nz2 = shape
    E0 = torch.from_numpy(smooth_field(Ny, Nz, 3)).to(device)
    d, z = 0.1e-6, 0.01
    L = MP.lens_struct({"focal_length": 500.0, "thickness": 1.0, "number": 1,
                        "absorption_sigma": 40.0}, LAM) if lens else None
    for mode in ("edge", "constant"):
        a = MP.asp_step(E0, d, d, LAM, z, Ny2, Nz2, padding_mode=mode, pad_constant=0.1 + 0.2j,
                        lens=L, method="pruned").cpu().numpy()
        b = MP.asp_step(E0, d, d, LAM, z, Ny2, Nz2, padding_mode=mode, pad_constant=0.1 + 0.2j,
                        lens=L, method="fft2d").cpu().numpy()
        e = rel(a, b)
        print(f"[parity] pruned vs padded {shape} {mode} lens={lens}: rel={e:.3e}")
        assert e <= TOL


def test_pruned_bench_size_vs_padded(device):
    """2048^2 field on 8192^2 pads (the CRL + 2 m workload's step), pruned
    against the padded hipFFT round trip."""
    n = 2048
    E0 = torch.from_numpy(smooth_field(n, n, 5)).to(device)
    d = 170e-9
    a = MP.asp_step(E0, d, d, LAM, 1.0, 8192, 8192, method="pruned")
    b = MP.asp_step(E0, d, d, LAM, 1.0, 8192, 8192, method="fft2d")
    e = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    print(f"[parity] pruned vs padded 2048^2 on 8192^2: rel={e:.3e}")
    assert e <= TOL


def test_evanescent_band(device):
    """Pitch below lambda/2: the outer spectrum decays."""
    E0 = smooth_field(64, 48, 1) + 0.3 * np.exp(1j * np.arange(64) * 2.9)[None, :].astype(np.complex64)
    d, z = 4.0e-11, 2.0e-10
    Ny2, Nz2 = MP.choose_optimal_pad(64, 48, d, d, LAM, z)
    got = MP.asp_step(torch.from_numpy(E0.astype(np.complex64)).to(device), d, d, LAM, z,
                      Ny2, Nz2).cpu().numpy()
    want = P.asp_step(E0, d, d, LAM, z, Ny2, Nz2)
    e = rel(got, want)
    print(f"[parity] evanescent: rel={e:.3e}")
    assert e <= TOL


def test_gaussian_lens_free_space_closed_form(device):
    """Gaussian through a thin lens (fused into the padding pass) and 30 mm of
    free space against the paraxial q-parameter law."""
    n, d, w0 = 256, 0.2e-6, 4e-6
    x = (np.arange(n) - (n - 1) / 2.0) * d
    K = 2 * np.pi / LAM

    def closed(f, z):
        a = 1.0 / w0 ** 2 + (1j * K / (2 * f) if f else 0.0)
        q = 1.0 + 2j * a * z / K
        return np.exp(-a * (x[None, :] ** 2 + x[:, None] ** 2) / q) / q

    ops = optics()
    ops.add_CRL_box(1, 200.0, 1.0)
    ops.add_free_space(30.0)
    b = beam()
    b.create_beam(10000.0)
    prop = b.free_space_propagator(padding_mode="constant")
    E = ops.apply_stack(closed(0.0, 0.0).astype(np.complex64), d, d, b.wavelength, prop)
    e = rel(E, closed(0.2, 0.03))
    print(f"[parity] gaussian lens+30mm closed form: rel={e:.3e}")
    assert e <= TOL


def test_wavefield_propagation_substeps_vs_oracle(device):
    """Detector field through CRL + free space with forced sub-steps, through
    beam.wavefield_propagation (the reference's entry point)."""
    det = detector()
    det.create_detector((192, 128), (1700.0, 1700.0))   # 170 nm pixels
    E0 = smooth_field(192, 128, 3)
    det.input_pixel_values(E0)
    ops = optics()
    ops.add_CRL_box(4, 500.0, 1.0, absorption_sigma=0.01, mu_per_m=300.0,
                    radius_of_curvature_m=2e-4)
    ops.add_free_space(40.0)
    b = beam()
    b.create_beam(10000.0)
    b.wavefield_propagation(det, ops, step_max=0.015)
    got = det.pixel_values
    want = P.apply_stack(ops.components, E0, 170e-9, 170e-9, LAM, step_max=0.015)
    e = rel(got, want)
    print(f"[parity] CRL + 40 mm in 3 steps: rel={e:.3e}")
    assert e <= TOL


def test_standalone_lens_vs_oracle(device):
    E0 = smooth_field(100, 90, 4)
    lens_d = {"kind": "lens box", "number": 2, "focal_length": 300.0, "thickness": 0.5,
              "absorption_sigma": 0.02, "mu_per_m": 100.0, "radius_of_curvature_m": 1e-4}
    L = MP.lens_struct(lens_d, LAM)
    t = torch.from_numpy(E0).to(device)
    got = MP.lens_apply(t, 170e-9, 150e-9, LAM, L).cpu().numpy()
    want = P.thin_lens(E0, 170e-9, 150e-9, lens_d, LAM)
    e = rel(got, want)
    print(f"[parity] lens: rel={e:.3e}")
    assert e <= TOL
