"""Detector: pixel geometry and the stored complex field.

Mirrors the parts of the reference ``detector`` class the kinematic path
consumes (Detector.py:73-231, 569-770):

* ``create_detector(shape, pixel_size, geometry="rectangular",
  construction_mode="plane", input_mode="spatial")`` builds pixel centres on
  the x = 0 plane at pitch ``pixel_size`` [Angstrom], centred on 0, float32,
  raveled with ``p = iz * Ny + iy`` (Detector.py:207-217);
* ``position_detector_absolute(distance, two_theta, eta, degrees=True)``
  rotates by R_eta(-x) @ R_2theta(-y) and translates to the float32 centre
  distance * (cos 2theta, sin 2theta sin eta, sin 2theta cos eta)
  (Detector.py:613-741);
* ``pixel_coordinates`` (3, Ny*Nz) float64 Angstrom, ``shape`` (Ny, Nz),
  ``pixel_size`` (dy, dz) Angstrom, ``input_pixel_values`` and the
  ``pixel_values`` / ``pixel_intensity`` views;
* ``add_noise`` (Poisson, Gaussian, uniform; Detector.py:772-846) and
  ``apply_psf`` (Gaussian, Lorentzian or pseudo-Voigt kernel, intensity or
  amplitude, scipy.ndimage boundary modes; Detector.py:848-1087), the
  convolution on the GPU.

Only rectangular spatial detectors in plane or shell construction are
implemented; ring and angular layouts raise ``NotImplementedError``.
"""
import numpy as np


class detector:
    def __init__(self, directory=None):
        self.directory = directory
        self._shape = None
        self._pixel_size = None
        self._pixel_coordinates = None
        self._pixel_values = None
        self._center = np.zeros(3)
        self._direction = np.array([1.0, 0.0, 0.0])
        self._two_theta = 0.0
        self._eta = 0.0
        self._distance = 0.0
        self._geometry = "rectangular"
        self._construction_mode = "plane"
        self._input_mode = "spatial"

    # ------------------------------------------------------------ geometry
    def create_detector(self, shape, pixel_size, geometry="rectangular",
                        construction_mode="plane", input_mode="spatial",
                        n_pixels=None):
        geometry = str(geometry).lower()
        construction_mode = str(construction_mode).lower()
        input_mode = str(input_mode).lower()
        if construction_mode not in ("plane", "shell"):
            raise ValueError(f"construction_mode must be 'plane' or 'shell', got '{construction_mode}'")
        if input_mode not in ("spatial", "angular"):
            raise ValueError(f"input_mode must be 'spatial' or 'angular', got '{input_mode}'")
        if geometry != "rectangular" or input_mode != "spatial":
            raise NotImplementedError("only rectangular spatial detectors are implemented")
        self._shape = (int(shape[0]), int(shape[1]))
        self._pixel_size = (float(pixel_size[0]), float(pixel_size[1]))
        self._geometry = geometry
        self._construction_mode = construction_mode
        self._input_mode = input_mode
        self._center = np.array([0.0, 0.0, 0.0])
        self._direction = np.array([1.0, 0.0, 0.0])
        self._two_theta = 0.0
        self._eta = 0.0
        self._distance = 0.0
        Ny0, Nz0 = self._shape
        y_lin = ((np.arange(Ny0) - (Ny0 - 1) / 2.0) * float(self._pixel_size[0])).astype(np.float32)
        z_lin = ((np.arange(Nz0) - (Nz0 - 1) / 2.0) * float(self._pixel_size[1])).astype(np.float32)
        Y, Z = np.meshgrid(y_lin, z_lin)
        X = np.full_like(Y, 0)
        self._pixel_coordinates = np.vstack((X.ravel(), Y.ravel(), Z.ravel()))

    @staticmethod
    def get_rotation(axis, angle):
        """Right-handed Rodrigues rotation by ``angle`` [rad] about ``axis``."""
        axis = np.asarray(axis, dtype=np.float64)
        axis = axis / np.linalg.norm(axis)
        c, s = np.cos(angle), np.sin(angle)
        d = 1.0 - c
        x, y, z = axis
        return np.array([[c + d * x * x, d * x * y - z * s, d * x * z + y * s],
                         [d * y * x + z * s, c + d * y * y, d * y * z - x * s],
                         [d * z * x - y * s, d * z * y + x * s, c + d * z * z]])

    def get_rotation_detector(self, two_theta, eta):
        eta_m = self.get_rotation(np.array([-1.0, 0.0, 0.0]), eta)
        tt_m = self.get_rotation(np.array([0.0, -1.0, 0.0]), two_theta)
        return eta_m @ tt_m

    def position_detector_absolute(self, distance, two_theta, eta, degrees=True):
        if self._shape is None:
            raise ValueError("Detector shape/pixel_size are not initialized. "
                             "Call create_detector(...) before positioning.")
        if degrees:
            two_theta = np.deg2rad(two_theta)
            eta = np.deg2rad(eta)
        self.create_detector(self._shape, self._pixel_size, geometry=self._geometry,
                             construction_mode=self._construction_mode,
                             input_mode=self._input_mode)
        self._two_theta = float(two_theta)
        self._eta = float(eta)
        self._distance = float(distance)
        R = self.get_rotation_detector(self._two_theta, self._eta)
        self._direction = R @ np.array([1.0, 0.0, 0.0], dtype=np.float32)
        self._center = self._distance * np.array([
            np.cos(self._two_theta),
            np.sin(self._two_theta) * np.sin(self._eta),
            np.sin(self._two_theta) * np.cos(self._eta)], dtype=np.float32)
        if self._construction_mode == "shell":
            pc = R @ self._pixel_coordinates + self._center[:, None]
            norms = np.linalg.norm(pc, axis=0, keepdims=True)
            self._pixel_coordinates = pc / norms * self._distance
        else:
            self._pixel_coordinates = R @ self._pixel_coordinates + self._center[:, None]

    def position_detector_relative(self, distance, two_theta, eta, degrees=True):
        dtt = np.deg2rad(two_theta) if degrees else float(two_theta)
        det = np.deg2rad(eta) if degrees else float(eta)
        self.position_detector_absolute(self._distance + float(distance),
                                        self._two_theta + dtt, self._eta + det,
                                        degrees=False)

    def get_detector_position_cartesian(self):
        return self._center, self._direction

    # ------------------------------------------------------------ field
    def input_pixel_values(self, pixel_values):
        self._pixel_values = pixel_values

    @property
    def pixel_values(self):
        return self._pixel_values

    @property
    def pixel_intensity(self):
        v = self._pixel_values
        return None if v is None else np.abs(v) ** 2

    @property
    def pixel_amplitude(self):
        v = self._pixel_values
        return None if v is None else np.abs(v)

    @property
    def pixel_phase(self):
        v = self._pixel_values
        return None if v is None else np.angle(v)

    # ------------------------------------------------ noise and resolution
    def add_noise(self, noise_type, photon_flux=None, sigma=None, seed=None):
        """Poisson, Gaussian or uniform noise on the stored field (add_noise,
        Detector.py:772-846).  The draws come from numpy's default_rng(seed)
        in the reference's order, so a seed gives the reference's realisation;
        the arithmetic around them is elementwise float32 (host)."""
        if self._pixel_values is None:
            raise ValueError("No pixel values to add noise to. Run a simulation first.")
        rng = np.random.default_rng(seed)
        E = np.array(self._pixel_values, dtype=np.complex64, copy=True)
        noise_type = str(noise_type).lower()
        if noise_type == "poisson":
            if photon_flux is None:
                raise ValueError("photon_flux is required for Poisson noise")
            photon_flux = float(photon_flux)
            intensity = np.abs(E) ** 2
            I_max = intensity.max()
            if I_max <= 0.0:
                return
            counts = rng.poisson(intensity * (photon_flux / I_max)).astype(np.float32)
            amp = np.sqrt(counts * (I_max / photon_flux))
            phase = np.angle(E)
            E = (amp * (np.cos(phase) + 1j * np.sin(phase))).astype(np.complex64)
        elif noise_type == "gaussian":
            if sigma is None:
                raise ValueError("sigma is required for Gaussian noise")
            scale = float(sigma) * np.abs(E).mean() / np.sqrt(2.0)
            noise = scale * (rng.standard_normal(E.shape).astype(np.float32)
                             + 1j * rng.standard_normal(E.shape).astype(np.float32))
            E = (E + noise).astype(np.complex64)
        elif noise_type == "uniform":
            if sigma is None:
                raise ValueError("sigma is required for uniform noise")
            scale = float(sigma) * np.abs(E).mean()
            noise = scale * (rng.uniform(-1, 1, E.shape).astype(np.float32)
                             + 1j * rng.uniform(-1, 1, E.shape).astype(np.float32))
            E = (E + noise).astype(np.complex64)
        else:
            raise ValueError(f"Unknown noise type '{noise_type}'. "
                             f"Use 'poisson', 'gaussian', or 'uniform'.")
        self.input_pixel_values(E)

    @staticmethod
    def _build_psf_kernel(sigma_ax0, sigma_ax1, profile="gaussian", rotation=0.0, eta_pv=0.5,
                          truncate=4.0):
        """Normalised PSF kernel on the pixel grid (float64), Gaussian,
        Lorentzian or pseudo-Voigt, rotated ellipse (Detector.py:848-871)."""
        max_sigma = max(sigma_ax0, sigma_ax1)
        if max_sigma <= 0:
            raise ValueError("PSF width must be positive")
        hw = max(1, int(np.ceil(truncate * max_sigma)))
        ax = np.arange(2 * hw + 1, dtype=np.float64) - hw
        g0, g1 = np.meshgrid(ax, ax, indexing="ij")
        rot = np.deg2rad(rotation)
        c, s = np.cos(rot), np.sin(rot)
        u = (c * g0 + s * g1) / sigma_ax0
        v = (-s * g0 + c * g1) / sigma_ax1
        r2 = u ** 2 + v ** 2
        profile = profile.lower()
        if profile == "gaussian":
            k = np.exp(-0.5 * r2)
        elif profile == "lorentzian":
            k = 1.0 / (1.0 + r2)
        elif profile == "pseudo-voigt":
            e = float(np.clip(eta_pv, 0.0, 1.0))
            k = (1.0 - e) * np.exp(-0.5 * r2) + e / (1.0 + r2)
        else:
            raise ValueError(f"Unknown profile '{profile}'. "
                             f"Use 'gaussian', 'lorentzian', or 'pseudo-voigt'.")
        tot = k.sum()
        if tot > 0:
            k /= tot
        return k

    _BOUNDARY = {"constant": 0, "reflect": 1, "nearest": 2, "mirror": 3, "wrap": 4}

    def apply_psf(self, width, profile="gaussian", rotation=0.0, units="pixels",
                  field="intensity", eta_pv=0.5, boundary="auto", truncate=4.0):
        """Convolve the detector intensity (or amplitude) with a PSF kernel,
        keeping the phase (apply_psf, Detector.py:873-1087); the convolution
        runs on the GPU (mosaic_psf_apply)."""
        import ctypes

        import torch

        from . import _lib
        if self._pixel_values is None:
            raise ValueError("No pixel values to apply PSF to. Run a simulation first.")
        if self._construction_mode == "shell" and self._input_mode != "angular":
            raise NotImplementedError(
                "PSF convolution is not supported for shell+spatial detectors because the pixel "
                "grid is non-regular after spherical projection. Use shell+angular mode instead.")
        if isinstance(width, (int, float, np.integer, np.floating)):
            s0 = s1 = float(width)
        elif hasattr(width, "__len__") and len(width) == 2:
            s0, s1 = float(width[0]), float(width[1])
        else:
            raise ValueError("width must be a scalar (circular) or 2-tuple (elliptical).")
        units = str(units).lower()
        if units == "physical":
            s0 /= float(self._pixel_size[0])
            s1 /= float(self._pixel_size[1])
        elif units == "angular":
            raise ValueError("units='angular' requires an angular-mode detector. "
                             "Use 'physical' or 'pixels'.")
        elif units != "pixels":
            raise ValueError(f"Unknown units '{units}'. Use 'pixels', 'physical', or 'angular'.")
        if s0 <= 0 or s1 <= 0:
            raise ValueError("PSF width must be positive after unit conversion.")
        kernel = np.ascontiguousarray(
            self._build_psf_kernel(s1, s0, profile=profile, rotation=rotation, eta_pv=eta_pv,
                                   truncate=truncate), dtype=np.float64)
        field = str(field).lower()
        if field not in ("intensity", "amplitude"):
            raise ValueError(f"Unknown field '{field}'. Use 'intensity' or 'amplitude'.")
        boundary = str(boundary).lower()
        mode = 0 if boundary == "auto" else self._BOUNDARY.get(boundary)
        if mode is None:
            raise ValueError(f"Unknown boundary '{boundary}'")
        rows, cols = self._shape[1], self._shape[0]
        dev = torch.device("cuda", torch.cuda.current_device())
        E = torch.as_tensor(np.ascontiguousarray(
            np.asarray(self._pixel_values, dtype=np.complex64).reshape(rows, cols)), device=dev)
        n = ctypes.c_size_t(0)
        _lib.check(_lib.lib().mosaic_psf_workspace(rows, cols, kernel.shape[0], ctypes.byref(n)),
                   "mosaic_psf_workspace")
        ws = torch.empty(max(1, n.value), dtype=torch.uint8, device=dev)
        _lib.check(_lib.lib().mosaic_psf_apply(
            ctypes.c_void_p(int(torch.view_as_real(E).data_ptr())), rows, cols,
            kernel.ctypes.data_as(ctypes.c_void_p), int(kernel.shape[0]), mode, mode,
            int(field == "amplitude"), ctypes.c_void_p(int(ws.data_ptr())), int(ws.numel()),
            ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))), "mosaic_psf_apply")
        self.input_pixel_values(E.cpu().numpy().ravel())

    @property
    def pixel_coordinates(self):
        return self._pixel_coordinates

    @property
    def shape(self):
        return self._shape

    @property
    def pixel_size(self):
        return self._pixel_size

    @property
    def distance(self):
        return self._distance

    @property
    def two_theta(self):
        return self._two_theta
