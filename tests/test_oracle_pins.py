"""Pins of the oracle's entrance-field sampler and defect field by what the
mathematics fixes, each written so that a plausible slip fails it.

Entrance field (the E0-only case of _ein_bilinear_cpu, Beam.py:956-1045):
bilinear interpolation reproduces any a + b i + c j + d i j exactly; at the
grid nodes it returns the node values; outside the grid it returns 0 (no edge
replication), while the last row and column are still inside.  With dyadic
values every one of these holds bit for bit in float32.  The same checks run
on the oracle with u and v swapped and with the centre index moved by one,
and must fail there.

Defect field (the device model of mosaic_defect_positions): a screw
dislocation gives u = b theta / 2 pi along the line on a circle about the
core; an edge dislocation's displacement jumps by its Burgers vector across
the cut (the Burgers circuit); the stacking fault shifts the atoms strictly
above its plane, as the reference's searchsorted(side="left") count of planes
below each atom does (Defects.py:3951-3990), and leaves an atom on the plane
in place.
"""
import numpy as np

from mosaic_amd import sample as S
from oracle import defects as OD
from oracle import kinematic as O

# ---------------------------------------------------------- entrance field
NYB, NZB = 12, 9
DU, DV, UC, VC = 2.0, 4.0, 5.5, 4.0


def _e0():
    i = np.arange(NYB, dtype=np.float64)[:, None]
    j = np.arange(NZB, dtype=np.float64)[None, :]
    re = 3.0 + 0.5 * i - 0.25 * j + 0.125 * i * j
    im = -1.0 + 0.25 * i + 0.5 * j - 0.0625 * i * j
    return (re + 1j * im).astype(np.complex64), (re, im)


def _exact(iu, iv):
    re = 3.0 + 0.5 * iu - 0.25 * iv + 0.125 * iu * iv
    im = -1.0 + 0.25 * iu + 0.5 * iv - 0.0625 * iu * iv
    return re + 1j * im


def _pos(iu, iv):
    """Lab positions whose beam coordinates (u = y, v = z) sit at fractional
    grid indices (iu, iv): y = (iu - uc) du, z = (iv - vc) dv."""
    iu = np.asarray(iu, dtype=np.float64)
    iv = np.asarray(iv, dtype=np.float64)
    p = np.zeros((iu.size, 3), dtype=np.float32)
    p[:, 0] = 17.0
    p[:, 1] = (iu - UC) * DU
    p[:, 2] = (iv - VC) * DV
    return p


def bilinear_failures(sampler):
    E0, _ = _e0()
    fails = []
    # off-node points with dyadic fractions: exact
    iu = np.array([0.25, 3.5, 7.75, 10.125, 0.5, 10.875])
    iv = np.array([0.5, 2.25, 7.75, 1.125, 7.875, 0.0625])
    got = sampler(_pos(iu, iv), E0, DU, DV, UC, VC)
    if not np.array_equal(got, _exact(iu, iv).astype(np.complex64)):
        fails.append("off-node")
    # nodes, including the last row and column
    ii, jj = np.meshgrid(np.arange(NYB), np.arange(NZB), indexing="ij")
    got = sampler(_pos(ii.ravel(), jj.ravel()), E0, DU, DV, UC, VC)
    if not np.array_equal(got, E0.ravel()):
        fails.append("nodes")
    # outside the grid: zero, also just past the last row / column
    iu = np.array([-0.25, NYB - 1 + 0.25, 3.0, 3.0, -3.0, NYB + 2.0])
    iv = np.array([2.0, 2.0, -0.125, NZB - 1 + 0.5, -1.0, NZB + 1.0])
    got = sampler(_pos(iu, iv), E0, DU, DV, UC, VC)
    if np.any(got != 0):
        fails.append("outside")
    return fails


def test_ein_oracle_pinned():
    assert bilinear_failures(O.ein_bilinear_numpy) == []


def test_ein_checks_catch_swapped_axes_and_shifted_centre():
    def swapped(pos, E0, du, dv, uc, vc):
        p = pos.copy()
        p[:, [1, 2]] = p[:, [2, 1]]
        return O.ein_bilinear_numpy(p, E0, du, dv, uc, vc)

    def shifted(pos, E0, du, dv, uc, vc):
        return O.ein_bilinear_numpy(pos, E0, du, dv, uc + 1.0, vc)

    def shifted_v(pos, E0, du, dv, uc, vc):
        return O.ein_bilinear_numpy(pos, E0, du, dv, uc, vc - 1.0)

    for mutant in (swapped, shifted, shifted_v):
        assert bilinear_failures(mutant), mutant.__name__


# ------------------------------------------------------------ defect field
def _params(core=(0.0, 0.0, 0.0), line=(0.0, 0.0, 1.0), e1=(1.0, 0.0, 0.0), bs=0.0, be=0.0,
            nu=0.3, eps=0.0, normal=(0.0, 0.0, 1.0), shift=(0.0, 0.0, 0.0), R=None, c0=1e9):
    line = np.asarray(line, float)
    e1 = np.asarray(e1, float)
    e2 = np.cross(line, e1)
    R = np.eye(3) if R is None else np.asarray(R, float)
    return np.concatenate([core, line, e1, e2, [bs, be, nu, eps], normal, shift, R.ravel(),
                           [c0]])


def _circle(radius, n, z=0.0, core=(0.0, 0.0)):
    th = np.linspace(-np.pi, np.pi, n, endpoint=False) + np.pi / (2 * n)
    return np.column_stack([core[0] + radius * np.cos(th), core[1] + radius * np.sin(th),
                            np.full(n, z)]), th


def test_screw_dislocation_closed_form():
    pts, th = _circle(25.0, 64, z=3.0, core=(2.0, -1.0))
    out = OD.defect_positions(pts, _params(core=(2.0, -1.0, 0.0), bs=4.0))
    u = out - pts
    assert np.allclose(u[:, :2], 0.0, atol=1e-12)
    assert np.allclose(u[:, 2], 4.0 * th / (2 * np.pi), atol=1e-12)
    # the same rule as the workload's screw field (u_z = b atan2(y - y0, x - x0) / 2 pi)
    assert np.allclose(u[:, 2], S.screw_dislocation_uz(pts, b=4.0, centre=(2.0, -1.0)),
                       atol=1e-12)


def test_edge_dislocation_burgers_circuit():
    """Along a closed circuit the displacement increments sum to zero except
    across the cut (theta = +-pi), where u jumps by the Burgers vector."""
    be = 2.5
    pts, _ = _circle(40.0, 4096)
    u = OD.defect_positions(pts, _params(be=be)) - pts
    du = np.diff(np.vstack([u, u[:1]]), axis=0)
    k = int(np.argmax(np.linalg.norm(du, axis=1)))
    # the jump less the smooth change over one step (the mean of its neighbours)
    jump = du[k] - 0.5 * (du[k - 1] + du[(k + 1) % len(du)])
    assert np.allclose(jump, [-be, 0.0, 0.0], atol=1e-6)     # theta wraps from +pi to -pi
    assert np.linalg.norm(np.sum(du, axis=0)) < 1e-9
    # far from the core the edge field's x part is b theta / 2 pi to O(1 / r^2) ...
    pts2, th2 = _circle(1e4, 16)
    u2 = OD.defect_positions(pts2, _params(be=be)) - pts2
    assert np.allclose(u2[:, 0], be * th2 / (2 * np.pi) + be / (2 * np.pi) * np.sin(2 * th2)
                       / (4 * (1 - 0.3)), atol=1e-6)


def test_stacking_fault_strictly_above_plane():
    rng = np.random.default_rng(0)
    pts = rng.uniform(-20, 20, size=(500, 3))
    n = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    c0 = float(pts[0] @ n)                   # the plane through atom 0: it stays
    shift = np.array([1.0, 1.0, -2.0]) * (4.08 / 6.0)
    out = OD.defect_positions(pts, _params(normal=n, shift=shift, c0=c0))
    above = pts @ n > c0
    assert np.allclose(out[above] - pts[above], shift, atol=1e-12)
    assert np.allclose(out[~above], pts[~above], atol=1e-12)
    assert not above[0]
    # the workload's fault rule is the same
    sf = S.stacking_fault_shift(pts, 4.08, plane_offset=c0)
    assert np.allclose(sf, out, atol=1e-12)


def test_strain_and_rotation():
    rng = np.random.default_rng(1)
    pts = rng.uniform(-10, 10, size=(50, 3))
    ang = 0.3
    R = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]])
    out = OD.defect_positions(pts, _params(eps=1e-3, R=R))
    assert np.allclose(out, (pts * (1 + 1e-3)) @ R.T, atol=1e-12)
