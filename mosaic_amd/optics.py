"""Optics stack: ordered free-space and CRL (thin lens) components.

Mirrors the reference ``optics`` class on the propagation path
(Optics.py:513-639, 743-793): ``add_free_space(length_mm)``,
``add_CRL_box(number, focal_length_mm, thickness_mm, absorption_sigma,
mu_per_m, radius_of_curvature_m)``, ``components``, and
``apply_stack(field, dx, dy, wavelength, propagate_free_space, use_gpu=True,
diagnostics=False) -> np.ndarray complex64``.

On the GPU the field stays resident between components.  When a lens box is
directly followed by free space and the callback is this build's propagator
(it carries ``accepts_lens``), the lens is fused into the first padding pass
of that propagation.  Other element kinds of the reference (Bragg magnifier,
angular filter, zone plate, aperture) are not part of this build and raise.
"""
import numpy as np
import torch

from . import propagation as P


class optics:
    def __init__(self, directory=None):
        self.directory = directory
        self._components = []

    @property
    def components(self):
        return self._components

    def add_free_space(self, length_mm):
        self._components.append({"kind": "free space", "length": float(length_mm)})

    def add_CRL_box(self, number, focal_length_mm, thickness_mm, absorption_sigma=np.inf,
                    mu_per_m=0.0, radius_of_curvature_m=float("inf")):
        self._components.append({
            "kind": "lens box", "number": int(number),
            "focal_length": float(focal_length_mm), "thickness": float(thickness_mm),
            "absorption_sigma": float(absorption_sigma), "mu_per_m": float(mu_per_m),
            "radius_of_curvature_m": float(radius_of_curvature_m)})

    @staticmethod
    def quadratic_phase_step(k_val, f, dx, dy, shape):
        """Largest per-pixel phase step of exp(-i k r^2 / 2f) at the field edge
        (_check_quadratic_phase_sampling, Optics.py:473-511)."""
        Ny, Nx = int(shape[0]), int(shape[1])
        f = abs(float(f))
        if f <= 0.0:
            return float("inf")
        return max(k_val * dx * ((Nx - 1) / 2.0 * dx) / f, k_val * dy * ((Ny - 1) / 2.0 * dy) / f)

    def apply_stack_device(self, field, dx, dy, wavelength, propagate_free_space):
        """apply_stack on a (Ny, Nx) complex64 device tensor; returns a device tensor."""
        E = field
        comps = self._components
        i = 0
        while i < len(comps):
            elem = comps[i]
            kind = str(elem.get("kind", "")).lower()
            if kind == "lens box":
                lens = P.lens_struct(elem, wavelength)
                nxt = comps[i + 1] if i + 1 < len(comps) else None
                if (nxt is not None and str(nxt.get("kind", "")).lower() == "free space"
                        and getattr(propagate_free_space, "accepts_lens", False)):
                    E = propagate_free_space(E, dx, dy, float(nxt.get("length", 0.0)) * 1e-3,
                                             lens=lens)
                    i += 2
                    continue
                E = P.lens_apply(E.contiguous().clone(), dx, dy, wavelength, lens)
            elif kind == "free space":
                E = propagate_free_space(E, dx, dy, float(elem.get("length", 0.0)) * 1e-3)
                if not isinstance(E, torch.Tensor):
                    E = torch.as_tensor(np.asarray(E, dtype=np.complex64), device=field.device)
            else:
                raise NotImplementedError(f'optics element "{kind}" is not part of this build')
            i += 1
        return E

    def apply_stack(self, field, dx, dy, wavelength, propagate_free_space, use_gpu=True,
                    diagnostics=False):
        if wavelength is None:
            raise ValueError("wavelength must be provided to optics.apply_stack")
        if not use_gpu:
            raise NotImplementedError("the CPU optics path is not part of this build")
        if not torch.cuda.is_available():
            raise RuntimeError("mosaic_amd runs on the GPU only: no HIP device is visible")
        dev = torch.device("cuda", torch.cuda.current_device())
        E = torch.as_tensor(np.ascontiguousarray(field, dtype=np.complex64), device=dev)
        E = self.apply_stack_device(E, float(dx), float(dy), float(wavelength),
                                    propagate_free_space)
        return E.cpu().numpy().astype(np.complex64)
