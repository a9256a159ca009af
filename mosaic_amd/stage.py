"""Goniometer stage: motors -> float32 sample-to-lab rotation and translation.

Mirrors Stage.py:74-92 (default goniometer), 252-270 (float32 Rodrigues),
322-382 (coupled axes, rotation composed on the left in motor order,
translation sum).  ``pos_lab = R @ pos_sample + T``; the beam applies
``pos @ R.T + T`` on row-stacked float32 positions (Beam.py:6279-6280).
"""
import numpy as np


class stage:
    def __init__(self, directory=None):
        self.directory = directory
        self.create_stage()

    def create_stage(self, name=("goniometer",),
                     motor_name=("mu", "phi", "chi", "omega", "x", "y", "z"),
                     motor_type=("R", "R", "R", "R", "T", "T", "T"),
                     motor_coupling=((None,), (0,), (0,), (0, 1, 2), (0, 1, 2, 3),
                                     (0, 1, 2, 3), (0, 1, 2, 3)),
                     motor_axis=((0.0, -1.0, 0.0), (0.0, -1.0, 0.0), (-1.0, 0.0, 0.0),
                                 (0.0, 0.0, -1.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                 (0.0, 0.0, 1.0)),
                     motor_value=(0.0,) * 7, motor_resolution=(0.0,) * 7):
        self._name = list(name)
        self._motor_name = np.array(motor_name)
        self._motor_type = np.array(motor_type)
        self._motor_coupling = [list(c) for c in motor_coupling]
        self._motor_axis = np.array(motor_axis, dtype=np.float64)
        self._motor_value = np.array(motor_value, dtype=np.float64)
        self._motor_resolution = np.array(motor_resolution, dtype=np.float64)
        self.get_rotation()
        self.get_translation()

    @staticmethod
    def get_axis_rotation(axis, angle):
        axis = np.asarray(axis, dtype=np.float64)
        axis = axis / np.linalg.norm(axis)
        c, s = np.cos(angle), np.sin(angle)
        d = 1.0 - c
        x, y, z = axis
        return np.array([[c + d * x * x, d * x * y - z * s, d * x * z + y * s],
                         [d * y * x + z * s, c + d * y * y, d * y * z - x * s],
                         [d * z * x - y * s, d * z * y + x * s, c + d * z * z]],
                        dtype=np.float32)

    def get_motor_index(self, name):
        idx = np.flatnonzero(self._motor_name == str(name))
        if idx.size == 0:
            raise ValueError(f"Unknown motor {name!r}")
        return int(idx[0])

    def get_motor_axis(self, motor_idx):
        axes = self._motor_axis.copy()
        for coupled in self._motor_coupling[motor_idx]:
            if coupled is None:
                break
            axes[motor_idx] = self.get_axis_rotation(
                self.get_motor_axis(coupled), self._motor_value[coupled]) @ axes[motor_idx]
        return axes[motor_idx]

    def get_rotation(self):
        R = np.eye(3, dtype=np.float32)
        for i in np.where(self._motor_type == "R")[0]:
            R = self.get_axis_rotation(self.get_motor_axis(i), self._motor_value[i]) @ R
        self._rotation = R
        return R

    def get_translation(self):
        T = np.zeros(3, dtype=np.float32)
        for i in np.where(self._motor_type == "T")[0]:
            T += self.get_motor_axis(i) * self._motor_value[i]
        self._translation = T
        return T

    def set_single_motor_value_absolute(self, name, value, degrees=True):
        i = self.get_motor_index(name)
        v = float(value)
        if self._motor_type[i] == "R" and degrees:
            v = float(np.deg2rad(v))
        self._motor_value[i] = v
        self.get_rotation()
        self.get_translation()

    def set_single_motor_value_relative(self, name, value, degrees=True):
        i = self.get_motor_index(name)
        v = float(value)
        if self._motor_type[i] == "R" and degrees:
            v = float(np.deg2rad(v))
        self._motor_value[i] += v
        self.get_rotation()
        self.get_translation()

    def get_single_motor_value(self, name, degrees=True):
        i = self.get_motor_index(name)
        v = float(self._motor_value[i])
        return float(np.rad2deg(v)) if (self._motor_type[i] == "R" and degrees) else v

    @property
    def rotation(self):
        return self._rotation

    @property
    def translation(self):
        return self._translation

    @property
    def motor_names(self):
        return list(self._motor_name)
