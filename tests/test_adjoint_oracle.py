"""The adjoint's oracle on the CPU: the analytic displacement gradient
sum Re(conj(g) i c_j) dphi/du_j equals central finite differences of
L = sum ((|A|^2 - I_t) / s)^2, atom by atom (h = 3e-4 A), with A from the long-double
kinematic restatement.  A plausible slip -- the sign of the phase gradient,
x_hat - u_hat reversed, a factor 2 pi or the Angstrom scale dropped -- moves
the two apart by O(1).
"""
import numpy as np

from mosaic_amd import tables
from mosaic_amd import workloads as W
from oracle import adjoint as AO
from oracle import kinematic as O


def _problem():
    w = W.fcc_fault_scan111(cells=(4, 4, 4), npix=12, n_angles=2, half_range_deg=0.05)
    b, d, st = w["beam"], w["detector"], w["stage"]
    pos = w["sample"].load_chunk_positions(1)
    rots, trs = [], []
    for a in w["angles_deg"]:
        st.set_single_motor_value_absolute("phi", float(a))
        rots.append(np.asarray(st.rotation, np.float32))
        trs.append(np.asarray(st.translation, np.float32))
    rots, trs = np.stack(rots), np.stack(trs)
    u = W.gaussian_bump_field(pos, 0.2, 4.0, np.random.default_rng(0))
    ut = W.gaussian_bump_field(pos, 0.2, 4.0, np.random.default_rng(1))
    k = 2 * np.pi / b.wavelength
    c = tables.species_constants(["Au"], b.energy)
    return w, pos, u, ut, rots, trs, k, c, b._polarization_state(), d.pixel_coordinates / 1e10


def _field(p, rots, trs, k, c, pol, pc, Ny, Nz):
    n = p.shape[0]
    out = []
    for R, T in zip(rots, trs):
        lab = (p.astype(np.float64) @ R.astype(np.float64).T + T.astype(np.float64)) / 1e10
        out.append(O.kin_sum_c(lab, np.repeat(c["wk"][0][None], n, 0), np.full(n, c["f0z"][0]),
                               np.full(n, c["fp"][0]), np.full(n, c["fpp"][0]), np.ones(n),
                               np.zeros(n), pc[0], pc[1], pc[2], k, False, Ny, Nz, True,
                               pol[0], pol[1:], False, precision="f80"))
    return np.stack(out)


def test_reference_gradient_matches_finite_differences():
    w, pos, u, ut, rots, trs, k, c, pol, pc = _problem()
    Ny, Nz = w["detector"].shape
    p_all = (pos + u).astype(np.float32)
    A = _field(p_all, rots, trs, k, c, pol, pc, Ny, Nz)
    It = np.abs(_field((pos + ut).astype(np.float32), rots, trs, k, c, pol, pc, Ny, Nz)) ** 2
    s = It.max()
    A, It = A / np.sqrt(s), It / s
    g = 4 * (np.abs(A) ** 2 - It) * A
    args = (c["wk"][0], float(c["f0z"][0]), float(c["fp"][0]), float(c["fpp"][0]),
            np.complex64(1.0), rots, trs, pc, k, pol)
    h = 3e-4            # truncation error ~h^2: 7e-6 here, 8e-5 at 1e-3
    ref, fd = [], []
    for j in np.random.default_rng(5).choice(pos.shape[0], 6, replace=False):
        cj = AO.atom_terms(p_all[j], *args) / np.sqrt(s)
        ref.append(AO.reference_gradient(cj, AO.phase_gradient(p_all[j], rots, trs, pc, k), g))
        row = []
        for ax in range(3):
            e = np.zeros(3)
            e[ax] = h
            cp = AO.atom_terms(p_all[j] + e, *args) / np.sqrt(s)
            cm = AO.atom_terms(p_all[j] - e, *args) / np.sqrt(s)
            row.append(AO.fd_gradient_intensity_loss(A, It, cj, cp, cm, h))
        fd.append(row)
    ref, fd = np.array(ref), np.array(fd)
    e = np.linalg.norm(ref - fd) / np.linalg.norm(ref)
    print(f"[oracle] adjoint reference vs FD: rel={e:.2e}")
    assert e < 1e-4
    # the gradient is not trivially small: moving one atom changes L visibly
    assert np.linalg.norm(ref) > 1e-6

