"""Oracle for the displacement gradient -- TEST INFRASTRUCTURE ONLY.

For a real loss L(A) of the kinematic field and the PyTorch gradient
convention g = dL/dRe(A) + i dL/dIm(A):

* ``atom_terms`` -- c_j(f, p), one atom's term of A for every pose and pixel
  (the C restatement of compute_scattering_cffi with one atom, long double by
  default; lab position R_f (s_j + u_j) + T_f evaluated in float64);
* ``reference_gradient`` -- dL/du_j = sum_{f,p} Re(conj(g) i c_j) dphi/du_j with
  the exact per-pair phase gradient dphi/du = k 1e-10 R^T (x_hat - u_hat)
  [per Angstrom], in float64;
* ``fd_gradient_intensity_loss`` -- central finite differences of
  L = sum (|A|^2 - I_t)^2, moving one atom at a time: A changes by that
  atom's term only, and L(+h) - L(-h) is formed as
  sum (I+ - I-)(I+ + I- - 2 I_t) to avoid cancellation.
"""
import numpy as np

from . import kinematic as O


def atom_terms(p_sample, f0p, f0z, fp, fpp, amp, rots, trans, pc, k, pol, threads=0,
               precision="f80"):
    """(F, npix) complex128 terms of one atom at sample-frame position p (Angstrom).

    The phase k (x + |T - r|) is ~5e10 rad at 1 m, so FP64 leaves ~1e-5 rad of
    rounding in every term: enough to swamp a finite difference over 1e-3 A
    (a 3e-3 rad change).  The default evaluates it in x87 long double.
    """
    F = rots.shape[0]
    out = np.zeros((F, pc.shape[1]), dtype=np.complex128)
    for f in range(F):
        lab = (rots[f].astype(np.float64) @ np.asarray(p_sample, dtype=np.float64)
               + trans[f].astype(np.float64)) / 1e10
        out[f] = O.kin_sum_c(lab[None, :], f0p[None, :], np.array([f0z]), np.array([fp]),
                             np.array([fpp]), np.array([amp.real]), np.array([amp.imag]),
                             pc[0], pc[1], pc[2], k, False, pc.shape[1], 1, True, pol[0],
                             pol[1:], False, precision=precision, threads=threads)
    return out


def phase_gradient(p_sample, rots, trans, pc, k):
    """(F, npix, 3) exact d phase / d u [rad per Angstrom], sample frame."""
    F = rots.shape[0]
    out = np.zeros((F, pc.shape[1], 3))
    for f in range(F):
        R = rots[f].astype(np.float64)
        lab = (R @ np.asarray(p_sample, dtype=np.float64) + trans[f].astype(np.float64)) / 1e10
        d = pc - lab[:, None]
        u = d / np.linalg.norm(d, axis=0)
        w = np.array([[1.0], [0.0], [0.0]]) - u
        out[f] = (k * 1e-10) * (R.T @ w).T
    return out


def reference_gradient(c_j, dphi, g):
    """sum Re(conj(g) i c_j) dphi over poses and pixels -> (3,)."""
    w = np.real(np.conj(g) * 1j * c_j)
    return np.einsum("fp,fpc->c", w, dphi)


def fd_gradient_intensity_loss(A, It, c0, c_plus, c_minus, h):
    """Central difference of sum (|A|^2 - It)^2 for one coordinate of one atom."""
    dp = c_plus - c0
    dm = c_minus - c0
    Ip = np.abs(A + dp) ** 2
    Im = np.abs(A + dm) ** 2
    diff = 2 * np.real(np.conj(A) * (dp - dm)) + np.abs(dp) ** 2 - np.abs(dm) ** 2
    return float(np.sum(diff * (Ip + Im - 2 * It)) / (2 * h))
