"""Oracle for the propagation path -- TEST INFRASTRUCTURE ONLY.

Restates, in NumPy float64 / complex128:

* the pad-size and sub-step rules ``_choose_optimal_pad`` /
  ``_alias_free_step`` (Beam.py:4632-4736) and the recursion of
  ``_angular_spectrum_propagate_cpu`` (Beam.py:7945-8022);
* one angular-spectrum step: pad (edge or constant), fft2, the transfer
  function of prop_mul_cpu (Beam.py:4853-4891) evaluated exactly in float64,
  ifft2, centre crop;
* the thin lens with CRL absorption ``_apply_thin_lens_box``
  (Optics.py:49-130) in float64;
* ``apply_stack`` for 'lens box' and 'free space' entries (Optics.py:513-639).

``ref_prop_mul`` calls the reference's own prop_mul_cpu built into
oracle/_ref by build_ref.py (float32 transfer function, in place).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_PROP = os.path.join(_HERE, "_ref", "libref_prop.so")


def next_pow_two(n):
    n_int = int(np.ceil(n))
    return 1 if n_int < 1 else 1 << (n_int - 1).bit_length()


def choose_optimal_pad(Ny, Nz, dy, dz, wavelength, z, safety=1.1, min_pad_factor=1.0):
    zabs = abs(float(z))
    if zabs == 0.0:
        return (next_pow_two(max(int(np.ceil(Ny * min_pad_factor)), Ny)),
                next_pow_two(max(int(np.ceil(Nz * min_pad_factor)), Nz)))
    out = []
    for n, d in ((Ny, dy), (Nz, dz)):
        sr = min(min(1.0, wavelength / (2.0 * d)), 0.999999)
        tan = min(sr / np.sqrt(max(1e-18, 1.0 - sr * sr)), 2.0)
        pad_px = int(np.ceil(safety * zabs * tan / d))
        n2 = max(n + 2 * pad_px, int(np.ceil(n * min_pad_factor)))
        out.append(next_pow_two(min(n2, 8192)))
    return tuple(out)


def alias_free_step(Ny, Nz, dy, dz, wavelength, z, pad_factor=1.0):
    Ny2, Nz2 = choose_optimal_pad(Ny, Nz, dy, dz, wavelength, z,
                                  min_pad_factor=max(1.0, pad_factor))
    zy, zz = Ny2 * dy ** 2 / wavelength, Nz2 * dz ** 2 / wavelength
    return min(zy, zz) if (zy > 0 and zz > 0) else float("inf")


def transfer_function(Ny2, Nz2, dy, dz, wavelength, z):
    """H on the unshifted fft2 grid (rows kz, columns ky), float64."""
    k = 2 * np.pi / wavelength
    ky = 2 * np.pi * np.fft.fftfreq(Ny2, d=dy)
    kz = 2 * np.pi * np.fft.fftfreq(Nz2, d=dz)
    return tf_from_k(ky, kz, k, z)


def tf_from_k(ky, kz, k, z):
    """H = exp(-i z kt^2 / (sqrt(k^2 - kt^2) + k)) where propagating, else
    exp(-|z| sqrt(kt^2 - k^2)) exp(-i z k) (prop_mul_cpu, Beam.py:4853-4891)."""
    ky = np.asarray(ky, dtype=np.float64)
    kz = np.asarray(kz, dtype=np.float64)
    k = float(k)
    z = float(z)
    kt2 = ky[None, :] ** 2 + kz[:, None] ** 2
    kx2 = k * k - kt2
    prop = kx2 >= 0
    ph = np.where(prop, -z * kt2 / (np.sqrt(np.abs(kx2)) + k), np.fmod(-z * k, 2 * np.pi))
    amp = np.where(prop, 1.0, np.exp(-abs(z) * np.sqrt(np.abs(kx2))))
    return amp * np.exp(1j * ph)


def asp_step(field, dy, dz, wavelength, z, Ny2, Nz2, padding_mode="edge", pad_constant=0.0):
    F0 = np.asarray(field, dtype=np.complex128)
    Nz, Ny = F0.shape
    z0, y0 = (Nz2 - Nz) // 2, (Ny2 - Ny) // 2
    if (padding_mode or "edge").lower() == "constant":
        Fp = np.full((Nz2, Ny2), complex(pad_constant), dtype=np.complex128)
        Fp[z0:z0 + Nz, y0:y0 + Ny] = F0
    else:
        Fp = np.pad(F0, ((z0, Nz2 - Nz - z0), (y0, Ny2 - Ny - y0)), mode="edge")
    Fp = np.fft.ifft2(np.fft.fft2(Fp) * transfer_function(Ny2, Nz2, dy, dz, wavelength, z))
    return Fp[z0:z0 + Nz, y0:y0 + Ny]


def propagate(field, dy, dz, wavelength, z, step_max=None, pad_factor=1.0,
              padding_mode="edge", pad_constant=0.0):
    F = np.asarray(field, dtype=np.complex128)
    Nz, Ny = F.shape
    limit = (alias_free_step(Ny, Nz, dy, dz, wavelength, z, pad_factor)
             if step_max is None else float(step_max))
    if abs(z) > limit:
        n = int(np.ceil(abs(z) / limit))
        for _ in range(n):
            F = propagate(F, dy, dz, wavelength, z / n, step_max=limit, pad_factor=pad_factor,
                          padding_mode=padding_mode, pad_constant=pad_constant)
        return F
    Ny2, Nz2 = choose_optimal_pad(Ny, Nz, dy, dz, wavelength, z,
                                  min_pad_factor=max(1.0, pad_factor))
    return asp_step(F, dy, dz, wavelength, z, Ny2, Nz2, padding_mode, pad_constant)


def thin_lens(field, dx, dy, lens_data, wavelength):
    k = 2 * np.pi / wavelength
    f = float(lens_data["focal_length"]) * 1e-3
    t = float(lens_data["thickness"]) * 1e-3
    nsig = float(lens_data.get("absorption_sigma", np.inf))
    n_l = int(lens_data["number"])
    mu = float(lens_data.get("mu_per_m", 0.0))
    r_ap = float(lens_data.get("radius_of_curvature_m", np.inf))
    E = np.asarray(field, dtype=np.complex128)
    Ny, Nx = E.shape
    x = (np.arange(Nx) - (Nx - 1) / 2.0) * dx
    y = (np.arange(Ny) - (Ny - 1) / 2.0) * dy
    R2 = x[None, :] ** 2 + y[:, None] ** 2
    E = E * np.exp(-1j * 0.5 * (k / f) * R2)
    if mu > 0.0 and np.isfinite(r_ap):
        E = E * np.exp(-(mu * n_l / (2.0 * r_ap)) * R2)
    if np.isfinite(nsig):
        E = E * np.exp(-n_l * t / nsig)
    return E


def apply_stack(components, field, dx, dy, wavelength, **prop_kw):
    E = np.asarray(field, dtype=np.complex128)
    for c in components:
        kind = str(c.get("kind", "")).lower()
        if kind == "free space":
            E = propagate(E, dx, dy, wavelength, float(c.get("length", 0.0)) * 1e-3, **prop_kw)
        elif kind == "lens box":
            E = thin_lens(E, dx, dy, c, wavelength)
        else:
            raise ValueError(kind)
    return E


def ref_prop_mul(F, ky, kz, k, z):
    """The reference's prop_mul_cpu on a complex64 spectrum (returns a copy)."""
    lib = ctypes.CDLL(REF_PROP)
    fn = lib.prop_mul_cpu
    fn.restype = None
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    out = np.ascontiguousarray(F, dtype=np.complex64).copy()
    ky = np.ascontiguousarray(ky, dtype=np.float32)
    kz = np.ascontiguousarray(kz, dtype=np.float32)
    Nz, Ny = out.shape
    fn(Ny, Nz, ky.ctypes.data, kz.ctypes.data, float(k), float(z), out.ctypes.data)
    return out
