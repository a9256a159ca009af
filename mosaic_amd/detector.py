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
  ``pixel_values`` / ``pixel_intensity`` views.

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
