"""NumPy restatement of the device defect field -- TEST INFRASTRUCTURE ONLY.

Same model as mosaic_defect_positions, in float64: a dislocation line through
``core`` along ``line`` with Burgers vector split into screw (b . line) and
edge parts; isotropic-elasticity displacement (Hirth & Lothe, Theory of
Dislocations, eqs. 3-45 and 3-46) in the frame (e1 along the edge component,
e2 = line x e1) with a 1 A^2 core regularisation of r^2; a hydrostatic
strain about the origin; a stacking-fault shift of the atoms on the positive
side of the plane n . r = c; then the crystal -> sample rotation.
"""
import numpy as np


def defect_positions(base, a):
    a = np.asarray(a, dtype=np.float64)
    core, line, e1, e2 = a[0:3], a[3:6], a[6:9], a[9:12]
    bs, be, nu, eps = a[12], a[13], a[14], a[15]
    n, shift, R, c0 = a[16:19], a[19:22], a[22:31].reshape(3, 3), a[31]
    r = np.asarray(base, dtype=np.float64)
    d = r - core
    x = d @ e1
    y = d @ e2
    r2 = x * x + y * y + 1.0
    th = np.arctan2(y, x)
    us = bs * th / (2 * np.pi)
    ux = be / (2 * np.pi) * (th + x * y / (2 * (1 - nu) * r2))
    uy = -be / (2 * np.pi) * ((1 - 2 * nu) / (4 * (1 - nu)) * np.log(r2)
                              + (x * x - y * y) / (4 * (1 - nu) * r2))
    u = us[:, None] * line + ux[:, None] * e1 + uy[:, None] * e2 + eps * r
    u[(r @ n) - c0 > 0] += shift
    return (r + u) @ R.T
