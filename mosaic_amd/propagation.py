"""Angular-spectrum propagation: host rules + the HIP step.

The pad-size and sub-step rules are the reference's host arithmetic,
restated:

* ``next_pow_two`` (Beam.py:4622-4629);
* ``choose_optimal_pad`` (Beam.py:4632-4701): half-padding
  |z| tan(theta_max) per axis with sin(theta_max) = min(1, lambda / 2d),
  times ``safety``, clamped to 8192, rounded up to a power of two;
* ``alias_free_step`` (Beam.py:4704-4736): z_max = N_pad d^2 / lambda;
* ``propagate`` (Beam.py:7829-7943): when |z| exceeds the step limit the
  distance is split into ceil(|z| / limit) equal sub-steps, each of which
  recomputes its pad size.

Each step runs in ``libmosaic_amd.so`` (``mosaic_asp_propagate``): by default
a pruned three-pass transform that never forms the padded grid (thin lens,
padding, transfer function and crop fused into its passes), else pad,
hipFFT forward, transfer function, inverse, crop.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import Lens, check

MAX_PADDED_DIM = 8192


def next_pow_two(n):
    n_int = int(np.ceil(n))
    if n_int < 1:
        return 1
    return 1 << (n_int - 1).bit_length()


def choose_optimal_pad(Ny, Nz, dy, dz, wavelength, z, safety=1.1, enforce_pow2=True,
                       min_pad_factor=1.0):
    zabs = abs(float(z))
    if zabs == 0.0:
        Ny2 = max(int(np.ceil(Ny * min_pad_factor)), Ny)
        Nz2 = max(int(np.ceil(Nz * min_pad_factor)), Nz)
        if enforce_pow2:
            Ny2, Nz2 = next_pow_two(Ny2), next_pow_two(Nz2)
        return int(Ny2), int(Nz2)
    sry = min(1.0, float(wavelength) / (2.0 * float(dy)))
    srz = min(1.0, float(wavelength) / (2.0 * float(dz)))
    sry = min(sry, 0.999999)
    srz = min(srz, 0.999999)
    tany = min(sry / np.sqrt(max(1e-18, 1.0 - sry * sry)), 2.0)
    tanz = min(srz / np.sqrt(max(1e-18, 1.0 - srz * srz)), 2.0)
    pad_y_px = int(np.ceil(safety * zabs * tany / float(dy)))
    pad_z_px = int(np.ceil(safety * zabs * tanz / float(dz)))
    Ny2 = max(Ny + 2 * pad_y_px, int(np.ceil(Ny * min_pad_factor)))
    Nz2 = max(Nz + 2 * pad_z_px, int(np.ceil(Nz * min_pad_factor)))
    Ny2 = min(Ny2, MAX_PADDED_DIM)
    Nz2 = min(Nz2, MAX_PADDED_DIM)
    if enforce_pow2:
        Ny2, Nz2 = next_pow_two(Ny2), next_pow_two(Nz2)
    return int(Ny2), int(Nz2)


def alias_free_step(Ny, Nz, dy, dz, wavelength, z, pad_factor=1.0):
    Ny2, Nz2 = choose_optimal_pad(int(Ny), int(Nz), float(dy), float(dz), float(wavelength),
                                  float(z), safety=1.1, enforce_pow2=True,
                                  min_pad_factor=max(1.0, float(pad_factor)))
    zy = Ny2 * float(dy) ** 2 / float(wavelength)
    zz = Nz2 * float(dz) ** 2 / float(wavelength)
    return float(min(zy, zz)) if (zy > 0 and zz > 0) else float("inf")


def plan_steps(Ny, Nz, dy, dz, wavelength, z, step_max=None, pad_factor=1.0):
    """The (z_step, Ny2, Nz2) sequence the reference's recursion would run."""
    z = float(z)
    limit = (alias_free_step(Ny, Nz, dy, dz, wavelength, z, pad_factor)
             if step_max is None else float(step_max))
    if abs(z) > limit:
        n = int(np.ceil(abs(z) / limit))
        steps = []
        for _ in range(n):
            steps += plan_steps(Ny, Nz, dy, dz, wavelength, z / n, step_max=limit,
                                pad_factor=pad_factor)
        return steps
    Ny2, Nz2 = choose_optimal_pad(Ny, Nz, float(dy), float(dz), float(wavelength), z,
                                  safety=1.1, enforce_pow2=True,
                                  min_pad_factor=max(1.0, float(pad_factor)))
    return [(z, Ny2, Nz2)]


def lens_struct(lens_data, wavelength):
    """mosaic_lens_t for an optics 'lens box' entry (Optics.py:49-130)."""
    f = float(lens_data["focal_length"]) * 1e-3
    t = float(lens_data["thickness"]) * 1e-3
    nsigma = float(lens_data.get("absorption_sigma", np.inf))
    n_l = int(lens_data["number"])
    mu = float(lens_data.get("mu_per_m", 0.0))
    r_apex = float(lens_data.get("radius_of_curvature_m", np.inf))
    L = Lens()
    L.focal_length_m = f
    L.parab_coef = (mu * n_l / (2.0 * r_apex)) if (mu > 0.0 and np.isfinite(r_apex)) else 0.0
    L.uniform_amp = float(np.float32(np.exp(-n_l * t / nsigma))) if np.isfinite(nsigma) else 1.0
    L.pad_ = 0.0
    return L


_WS = {}


def _workspace(Ny2, Nz2, device):
    out = ctypes.c_size_t(0)
    check(_lib.lib().mosaic_asp_workspace(int(Ny2), int(Nz2), ctypes.byref(out)),
          "mosaic_asp_workspace")
    need = int(out.value)
    key = (device.index, need)
    ws = _WS.get(device.index)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=device)
        _WS[device.index] = ws
    return ws


def _ptr(t):
    return ctypes.c_void_p(int(t.data_ptr()))


METHODS = {"auto": 0, "fft2d": 1, "pruned": 2}


def asp_step(field, dy, dz, wavelength, z, Ny2, Nz2, padding_mode="edge", pad_constant=0.0,
             lens=None, out=None, stream=None, method="auto"):
    """One step on a (Nz, Ny) complex64 device tensor; returns the new field.

    ``method``: "auto" (the pruned three-pass transform where the sizes allow
    it), "fft2d" (padded 2-D hipFFT round trip) or "pruned"."""
    if field.dtype != torch.complex64 or field.dim() != 2 or not field.is_cuda:
        raise ValueError("field must be a (Nz, Ny) complex64 device tensor")
    field = field.contiguous()
    Nz, Ny = int(field.shape[0]), int(field.shape[1])
    if out is None:
        out = torch.empty_like(field)
    ws = _workspace(Ny2, Nz2, field.device)
    pc = complex(pad_constant)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(_lib.lib().mosaic_asp_propagate(
        _ptr(torch.view_as_real(field)), Ny, Nz, float(dy), float(dz), float(wavelength),
        float(z), int(Ny2), int(Nz2), int((padding_mode or "edge").lower() != "constant"),
        float(pc.real), float(pc.imag), ctypes.byref(lens) if lens is not None else None,
        METHODS[method], _ptr(torch.view_as_real(out)), _ptr(ws), int(ws.numel()),
        ctypes.c_void_p(int(s.cuda_stream))), "mosaic_asp_propagate")
    return out


def propagate(field, dy, dz, wavelength, z, step_max=None, pad_factor=1.0, padding_mode="edge",
              pad_constant=0.0, lens=None, method="auto"):
    """Free-space propagation with the reference's automatic sub-stepping.

    ``lens`` (a mosaic_lens_t) is applied to the input field before the
    first step.
    """
    Nz, Ny = int(field.shape[0]), int(field.shape[1])
    E = field
    for i, (zs, Ny2, Nz2) in enumerate(plan_steps(Ny, Nz, dy, dz, wavelength, z, step_max,
                                                   pad_factor)):
        E = asp_step(E, dy, dz, wavelength, zs, Ny2, Nz2, padding_mode, pad_constant,
                     lens=lens if i == 0 else None, method=method)
    return E


def lens_apply(field, dx, dy, wavelength, lens):
    """Thin lens in place on a (Nz, Ny) complex64 device tensor."""
    if field.dtype != torch.complex64 or not field.is_contiguous():
        raise ValueError("field must be a contiguous complex64 device tensor")
    Nz, Ny = int(field.shape[0]), int(field.shape[1])
    s = torch.cuda.current_stream()
    check(_lib.lib().mosaic_lens_apply(_ptr(torch.view_as_real(field)), Ny, Nz, float(dx),
                                       float(dy), float(wavelength), ctypes.byref(lens),
                                       ctypes.c_void_p(int(s.cuda_stream))),
          "mosaic_lens_apply")
    return field
