"""Oracle for the depth-dependent entrance field -- TEST INFRASTRUCTURE ONLY.

NumPy restatement, in float32 and the reference's operation order, of
* the depth window: min and max of the stage-transformed positions along
  the +x beam (_compute_global_depth_bounds, Beam.py:5583-5620);
* the beam-column transmission map A(u, v) = exp(-tau + i phi): a 3 x 3
  triangular-shaped-cloud deposit of each in-grid atom's forward factors
  (f0(0) + f', f''), phi = -k C sum_re, tau = max(k C sum_im, 0) with
  C = r_e lambda^2 / (2 pi du dv), an optional Gaussian blur as an FFT
  product (_compute_beam_column_A_map_cpu, Beam.py:7118-7282);
* the sampler: tau' = -ln(|A| + 1e-20), phi' = arg A, E0, tau', phi'
  bilinear at (u, v), f = clip((x - s_min) / (s_max - s_min), 0, 1),
  Ein = E0 exp(-f tau') exp(i f phi') (_ein_bilinear_cpu, Beam.py:956-1045;
  precompute_depth_ein_all_chunks, Beam.py:5044-5060);
* the M-channel incident amplitude sum_m E_m e^{2 pi i g_m . r}
  (Beam.py:2283-2309).
"""
import numpy as np


def lab_positions(pos, R, T):
    """(pos @ R.T) + T in float32."""
    return (np.asarray(pos, np.float32) @ np.asarray(R, np.float32).T
            + np.asarray(T, np.float32)).astype(np.float32)


def depth_bounds(pos_lab):
    s = np.asarray(pos_lab, np.float32)[:, 0]
    if s.size == 0:
        return 0.0, 1.0
    lo, hi = np.float32(np.min(s)), np.float32(np.max(s))
    if not np.isfinite(lo) or hi <= lo:
        return 0.0, 1.0
    return float(lo), float(hi)


def _tsc_w(d):
    w = np.zeros_like(d, dtype=np.float32)
    m0 = d <= 0.5
    w[m0] = 0.75 - d[m0] * d[m0]
    m1 = (~m0) & (d <= 1.5)
    t = 1.5 - d[m1]
    w[m1] = 0.5 * t * t
    return w


def column_sums(pos_lab, fr, fi, NyB, NzB, du, dv, uc, vc):
    """TSC deposit of per-atom forward factors (fr, fi) onto the beam grid."""
    sum_re = np.zeros((NyB, NzB), np.float32)
    sum_im = np.zeros((NyB, NzB), np.float32)
    p = np.asarray(pos_lab, np.float32)
    iu = p[:, 1] / np.float32(du) + np.float32(uc)
    iv = p[:, 2] / np.float32(dv) + np.float32(vc)
    inb = (iu >= 0.0) & (iu <= NyB - 1) & (iv >= 0.0) & (iv <= NzB - 1)
    iu, iv = iu[inb], iv[inb]
    fr = np.asarray(fr, np.float32)[inb]
    fi = np.asarray(fi, np.float32)[inb]
    ic = np.floor(iu + np.float32(0.5)).astype(np.int64)
    jc = np.floor(iv + np.float32(0.5)).astype(np.int64)
    for dx in (-1, 0, 1):
        wx = _tsc_w(np.abs(iu - (ic + dx).astype(np.float32)))
        ii = ic + dx
        for dy in (-1, 0, 1):
            wy = _tsc_w(np.abs(iv - (jc + dy).astype(np.float32)))
            jj = jc + dy
            fac = wx * wy
            m = (ii >= 0) & (ii < NyB) & (jj >= 0) & (jj < NzB) & (fac > 0.0)
            idx = ii[m] * NzB + jj[m]
            np.add.at(sum_re.ravel(), idx, (fr[m] * fac[m]).astype(np.float32))
            np.add.at(sum_im.ravel(), idx, (fi[m] * fac[m]).astype(np.float32))
    return sum_re, sum_im


def _blur(a, radius):
    rad = int(radius)
    sig = rad / 2.0
    y, x = np.ogrid[-rad:rad + 1, -rad:rad + 1]
    k = np.exp(-(x * x + y * y) / (2.0 * sig * sig)).astype(np.float32)
    k /= k.sum()
    Fk = np.fft.fft2(k, s=a.shape)
    return np.fft.ifft2(np.fft.fft2(a) * Fk).real.astype(np.float32)


def column_map(sum_re, sum_im, du, dv, wavelength_A, kernel_radius=0):
    """A(u, v) complex64 from the column sums."""
    r_e_A = 2.81794092e-5
    C = (r_e_A * (wavelength_A * wavelength_A)) / (2.0 * np.pi * (du * dv))
    kA = 2.0 * np.pi / wavelength_A
    phi = np.float32(-kA) * (np.float32(C) * sum_re.astype(np.float32))
    tau = np.float32(kA) * (np.float32(C) * sum_im.astype(np.float32))
    tau = np.maximum(tau, np.float32(0.0))
    if kernel_radius > 0:
        phi = _blur(phi, kernel_radius)
        tau = np.maximum(_blur(tau, kernel_radius), np.float32(0.0))
    return np.exp(-tau + 1j * phi).astype(np.complex64)


def sampled_maps(A):
    tau = (-np.log(np.abs(A) + np.float32(1e-20))).astype(np.float32)
    phi = np.angle(A).astype(np.float32)
    return tau, phi


def ein_depth(pos_lab, tau, phi, E0, du, dv, uc, vc, s_min, s_max):
    """Per-atom Ein (complex64); zero outside the grid."""
    p = np.asarray(pos_lab, np.float32)
    N = p.shape[0]
    out = np.zeros(N, np.complex64)
    NyB, NzB = tau.shape
    iu = p[:, 1] / np.float32(du) + np.float32(uc)
    iv = p[:, 2] / np.float32(dv) + np.float32(vc)
    inb = (iu >= 0.0) & (iu <= NyB - 1) & (iv >= 0.0) & (iv <= NzB - 1)
    if not np.any(inb):
        return out
    iu, iv = iu[inb], iv[inb]
    i0 = np.floor(iu).astype(np.int64)
    j0 = np.floor(iv).astype(np.int64)
    i1 = np.clip(i0 + 1, 0, NyB - 1)
    j1 = np.clip(j0 + 1, 0, NzB - 1)
    fu = (iu - i0).astype(np.float32)
    fv = (iv - j0).astype(np.float32)
    w00 = (1.0 - fu) * (1.0 - fv)
    w01 = (1.0 - fu) * fv
    w10 = fu * (1.0 - fv)
    w11 = fu * fv
    r00, r01, r10, r11 = i0 * NzB + j0, i0 * NzB + j1, i1 * NzB + j0, i1 * NzB + j1
    t, ph, e = tau.ravel(), phi.ravel(), np.asarray(E0, np.complex64).ravel()
    ts = t[r00] * w00 + t[r01] * w01 + t[r10] * w10 + t[r11] * w11
    ps = ph[r00] * w00 + ph[r01] * w01 + ph[r10] * w10 + ph[r11] * w11
    es = e[r00] * w00 + e[r01] * w01 + e[r10] * w10 + e[r11] * w11
    denom = float(s_max) - float(s_min)
    if not denom > 0.0:
        denom = 1.0
    f = np.clip((p[inb, 0] - float(s_min)) / denom, 0.0, 1.0).astype(np.float32)
    amp = np.exp(-f * ts).astype(np.float32)
    phase = f * ps
    c, s = np.cos(phase), np.sin(phase)
    out[inb] = ((es.real * c - es.imag * s) * amp + 1j * (es.real * s + es.imag * c) * amp
                ).astype(np.complex64)
    return out


def mbeam_amplitude(pos_lab_m, E, g):
    """M-channel incident amplitude sum_m E[:, m] exp(2 pi i g_m . r) in
    float64, r the lab position in metres and g (M, 3) in cycles per metre
    (the lattice-phase sum of the general kernel, Beam.py:2283-2309; the
    reference forms g . r with an error-free split in float32 and reduces it
    modulo one cycle, which float64 does directly at these magnitudes)."""
    r = np.asarray(pos_lab_m, np.float64)
    G = np.asarray(g, np.float64).reshape(-1, 3)
    t = r @ G.T
    t -= np.rint(t)
    return np.sum(np.asarray(E, np.complex128) * np.exp(2j * np.pi * t), axis=1)
