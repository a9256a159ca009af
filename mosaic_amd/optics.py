"""Optics stack: ordered free-space, CRL (thin lens), Bragg magnifier,
angular filter, aperture and Fresnel zone plate components.

Mirrors the reference ``optics`` class on the propagation path
(Optics.py:132-471, 513-639, 743-957): the ``add_*`` methods append the same
component dicts, ``components``, and ``apply_stack(field, dx, dy,
wavelength, propagate_free_space, use_gpu=True, diagnostics=False) ->
np.ndarray complex64``.

On the GPU the field stays resident between components, and every element is
a HIP kernel of libmosaic_amd.so (propagate.hip, optics.hip).  When a lens box
is directly followed by free space and the callback is this build's
propagator (it carries ``accepts_lens``), the lens is fused into the first
padding pass of that propagation.  An unknown component kind raises
``ValueError`` as the reference does.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import propagation as P
from ._lib import AngularFilter, ZonePlate, check


def _ptr(t):
    return ctypes.c_void_p(int(t.data_ptr()))


def _stream():
    return ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))


def _f32(E):
    return torch.view_as_real(E.contiguous())


def bragg_magnifier(E, elem):
    """Two-bounce Bragg magnifier on a (Ny, Nx) complex64 device field."""
    E = E.contiguous()
    out = torch.empty_like(E)
    Ny, Nx = E.shape
    check(_lib.lib().mosaic_bragg_magnifier(
        _ptr(_f32(E)), _ptr(_f32(out)), int(Ny), int(Nx),
        float(elem.get("magnification_x", 1.0)), float(elem.get("magnification_y", 1.0)),
        float(elem.get("reflectivity", 1.0)), float(elem.get("phase_shift", 0.0)),
        int(elem.get("order", 1)), int(str(elem.get("pad_mode", "zeros")).lower() == "edge"),
        int(bool(elem.get("conserve_energy", True))), _stream()), "mosaic_bragg_magnifier")
    return out


_ROLLOFF = {"tophat": 0, "darwin": 1}


def angular_filter(E, dx, dy, filt, wavelength):
    """k-space angular filter, in place on a copy of the (Ny, Nx) field."""
    E = E.contiguous().clone()
    Ny, Nx = E.shape
    f = AngularFilter()
    f.center_x_mrad = float(filt.get("center_x_mrad", 0.0))
    f.center_y_mrad = float(filt.get("center_y_mrad", 0.0))
    hx = float(filt.get("half_angle_x_mrad", 0.0))
    f.half_angle_x_mrad = hx
    f.half_angle_y_mrad = float(filt.get("half_angle_y_mrad", hx))
    f.transmission = float(filt.get("transmission", 1.0))
    f.phase_shift = float(filt.get("phase_shift", 0.0))
    f.roll_deg = float(filt.get("roll_deg", 0.0))
    f.mode_1d = int(str(filt.get("mode", filt.get("dimension", "2d"))).lower() == "1d")
    f.elliptical = int(str(filt.get("shape", "circular")).lower() == "elliptical")
    f.rolloff = _ROLLOFF.get(str(filt.get("rolloff", "tophat")).lower(), 2)
    f.order = max(1, int(filt.get("order", 4)))
    n = ctypes.c_size_t(0)
    check(_lib.lib().mosaic_angular_filter_workspace(int(Ny), int(Nx), ctypes.byref(n)),
          "mosaic_angular_filter_workspace")
    ws = torch.empty(max(1, n.value), dtype=torch.uint8, device=E.device)
    check(_lib.lib().mosaic_angular_filter(_ptr(_f32(E)), int(Ny), int(Nx), float(dx), float(dy),
                                           float(wavelength), ctypes.byref(f), _ptr(ws),
                                           int(ws.numel()), _stream()), "mosaic_angular_filter")
    return E


_ZP_TYPES = {"amplitude": 0, "phase": 1, "ideal": 2}


def zone_plate(E, dx, dy, zp, wavelength):
    E = E.contiguous().clone()
    Ny, Nx = E.shape
    z = ZonePlate()
    z.outermost_zone_width_nm = float(zp["outermost_zone_width_nm"])
    z.diameter_um = float(zp["diameter_um"])
    z.efficiency = float(zp.get("efficiency", 1.0))
    z.central_stop_fraction = float(zp.get("central_stop_fraction", 0.0))
    z.edge_taper = float(zp.get("edge_taper", 0.0))
    z.order = max(1, int(zp.get("order", 1)))
    z.type = _ZP_TYPES.get(str(zp.get("zone_plate_type", "amplitude")).lower(), 0)
    check(_lib.lib().mosaic_zone_plate(_ptr(_f32(E)), int(Ny), int(Nx), float(dx), float(dy),
                                       float(wavelength), ctypes.byref(z), _stream()),
          "mosaic_zone_plate")
    return E


def aperture(E, dx, dy, ap):
    E = E.contiguous().clone()
    Ny, Nx = E.shape
    shape = str(ap.get("shape", ap.get("type", "square"))).lower()
    check(_lib.lib().mosaic_aperture(_ptr(_f32(E)), int(Ny), int(Nx), float(dx), float(dy),
                                     float(ap["width"]), int(shape == "circular"), _stream()),
          "mosaic_aperture")
    return E


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

    def add_bragg_magnifier_2b(self, magnification_x, magnification_y, reflectivity=1.0,
                               phase_shift=0.0, order=1, pad_mode="zeros",
                               conserve_energy=True):
        self._components.append({
            "kind": "bragg magnifier 2b", "magnification_x": float(magnification_x),
            "magnification_y": float(magnification_y), "reflectivity": float(reflectivity),
            "phase_shift": float(phase_shift), "order": int(order), "pad_mode": str(pad_mode),
            "conserve_energy": bool(conserve_energy)})

    def add_angular_filter(self, half_angle_mrad, center_x_mrad=0.0, center_y_mrad=0.0,
                           shape="circular", half_angle_y_mrad=None, rolloff="tophat", order=4,
                           transmission=1.0, phase_shift=0.0, mode="2d", roll_deg=0.0):
        if shape.lower() not in ("circular", "elliptical"):
            shape = "circular"
        if half_angle_y_mrad is None:
            half_angle_y_mrad = half_angle_mrad
        self._components.append({
            "kind": "angular filter", "shape": shape.lower(),
            "center_x_mrad": float(center_x_mrad), "center_y_mrad": float(center_y_mrad),
            "half_angle_x_mrad": float(half_angle_mrad),
            "half_angle_y_mrad": float(half_angle_y_mrad), "rolloff": str(rolloff).lower(),
            "order": int(order), "transmission": float(transmission),
            "phase_shift": float(phase_shift), "mode": str(mode).lower(),
            "roll_deg": float(roll_deg)})

    def add_aperture(self, width_mm, shape="square"):
        self._components.append({"kind": "aperture", "type": shape.lower(),
                                 "width": float(width_mm)})

    def add_fresnel_zone_plate(self, outermost_zone_width_nm, diameter_um, efficiency=1.0,
                               order=1, central_stop_fraction=0.0, zone_plate_type="amplitude",
                               edge_taper=0.0):
        self._components.append({
            "kind": "zone plate", "outermost_zone_width_nm": float(outermost_zone_width_nm),
            "diameter_um": float(diameter_um), "efficiency": float(efficiency),
            "order": int(order), "central_stop_fraction": float(central_stop_fraction),
            "zone_plate_type": str(zone_plate_type).lower(), "edge_taper": float(edge_taper)})

    def add_custom_component(self, component):
        self._components.append(component)

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
            elif kind == "bragg magnifier 2b":
                E = bragg_magnifier(E, elem)
            elif kind == "angular filter":
                E = angular_filter(E, dx, dy, elem, wavelength)
            elif kind == "aperture":
                E = aperture(E, dx, dy, elem)
            elif kind == "zone plate":
                E = zone_plate(E, dx, dy, elem, wavelength)
            else:
                raise ValueError(f'Unknown optics element "{kind}"')
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
