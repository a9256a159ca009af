"""Oracle for the kinematic sum -- TEST INFRASTRUCTURE ONLY (see package doc).

``kin_sum_numpy`` restates compute_scattering_cffi (Beam.py:663-816) and its
NumPy twin ``_cpu_scatter_numpy`` (Beam.py:848-953) in vectorised NumPy, in
float64 (the reference's precision) or ``np.longdouble`` (80-bit on x86, the
high-precision truth).  ``kin_sum_c`` calls the C restatement in
``kin_sum.c`` (OpenMP over pixels) and ``ref_kin_sum`` calls the reference's
own C kernel compiled into ``oracle/_ref`` by ``build_ref.py``.

All three take the same arguments as the reference kernel: positions (N, 3)
float64 metres, per-atom WK rows (N, 11) float32, f0(0), f', f'', incident
amplitude re/im (float32), pixel coordinates (Ny*Nz,) float64 metres each,
k [rad/m], polarization (p, e), flags.  They return complex (Ny*Nz,) sums
(not reshaped); the reference's callers reshape to (Nz, Ny).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_kin.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libref_scatter.so")


def q_cut(cx, cy, cz, Ny, Nz, k_val, dtype=np.float64):
    """Forward-removal cut-off per pixel (Beam.py:718-746)."""
    cx = np.asarray(cx, dtype=dtype)
    cy = np.asarray(cy, dtype=dtype)
    cz = np.asarray(cz, dtype=dtype)
    P = Ny * Nz
    R0 = np.sqrt(cx * cx + cy * cy + cz * cz)
    safe = np.where(R0 > 0, R0, 1)
    u = np.stack([cx, cy, cz], axis=1) / safe[:, None]
    u[R0 <= 0] = 0
    p = np.arange(P)
    ix, iy = p % Ny, p // Ny
    n_right = np.where(ix + 1 < Ny, p + 1, np.where(ix > 0, p - 1, p))
    n_up = np.where(iy + 1 < Nz, p + Ny, np.where(iy > 0, p - Ny, p))

    def step(nb):
        rr = np.stack([cx[nb], cy[nb], cz[nb]], axis=1)
        Rr = np.sqrt((rr * rr).sum(axis=1))
        cosd = np.where(Rr > 0, (u * rr).sum(axis=1) / np.where(Rr > 0, Rr, 1), 1)
        cosd = np.clip(cosd, -1, 1)
        return dtype(k_val) * np.sqrt(np.maximum(0, 2 * (1 - cosd)))

    qx, qy = step(n_right), step(n_up)
    return 0.5 * np.sqrt(qx * qx + qy * qy)


def analyser_weight(kind, d_unit, axis, accept_rad, darwin_hw_rad):
    """Analyser acceptance of the reference's general kernel (Beam.py:2350-2378)
    for unit rays ``d_unit`` (..., 3) from atom to pixel: 1 inside the cone of
    half-angle ``accept_rad`` about ``axis`` (top-hat, kind 1), or the Darwin
    roll-off 1 / (1 + (delta / hw)^2) (kind 2); delta the ray's angle to the
    axis, here atan2(|c x d|, c . d), the intended rule (the reference compares
    float32 cosines, which rounds cos(accept) to 1 below ~2.4e-4 rad)."""
    c = np.asarray(axis, dtype=np.float64)
    c = c / np.linalg.norm(c)
    cr = np.cross(np.broadcast_to(c, d_unit.shape), d_unit)
    delta = np.arctan2(np.sqrt((cr * cr).sum(axis=-1)), (d_unit * c).sum(axis=-1))
    if kind == 1:
        return (delta <= accept_rad).astype(d_unit.dtype)
    if kind == 2 and darwin_hw_rad > 0:
        r = delta / darwin_hw_rad
        return 1.0 / (1.0 + r * r)
    return np.ones(delta.shape, dtype=d_unit.dtype)


def kin_sum_numpy(pos_m, f0p, f0z, sanr, sani, amp_r, amp_i, cx, cy, cz, k_val,
                  remove_forward, Ny, Nz, apply_pol, pol_deg, pol_e, apply_decay,
                  dtype=np.float64, tile=256, analyser=None):
    """Vectorised restatement of the reference kernel; returns complex sums.

    ``analyser`` = (kind, axis, accept_rad, darwin_hw_rad) adds the general
    kernel's analyser weight per pair (see ``analyser_weight``)."""
    dt = dtype
    two_pi = dt(2) * (dt(np.pi) if dt is np.float64 else np.arccos(dt(-1)))
    pi_r = two_pi / 2
    cx = np.asarray(cx, dtype=dt)
    cy = np.asarray(cy, dtype=dt)
    cz = np.asarray(cz, dtype=dt)
    k = dt(k_val)
    R0 = np.sqrt(cx * cx + cy * cy + cz * cz)
    qc = q_cut(cx, cy, cz, Ny, Nz, k, dtype=dt)
    pos = np.asarray(pos_m, dtype=dt)
    f0p = np.asarray(f0p, dtype=dt)
    a_, c_, b_ = f0p[:, 0:5], f0p[:, 5], f0p[:, 6:11]
    f0z = np.asarray(f0z, dtype=dt)
    sanr = np.asarray(sanr, dtype=dt)
    sani = np.asarray(sani, dtype=dt)
    amp_r = np.asarray(amp_r, dtype=dt)
    amp_i = np.asarray(amp_i, dtype=dt)
    ex, ey, ez = (dt(v) for v in pol_e)
    pdeg = dt(pol_deg)
    s_fac = dt(0.25e-10) / pi_r
    rE = dt(2.81794092e-15)
    acc_r = np.zeros(cx.shape[0], dtype=dt)
    acc_i = np.zeros(cx.shape[0], dtype=dt)
    n = pos.shape[0]
    for a0 in range(0, n, tile):
        a1 = min(a0 + tile, n)
        ax = pos[a0:a1, 0][:, None]
        ay = pos[a0:a1, 1][:, None]
        az = pos[a0:a1, 2][:, None]
        dx = cx[None, :] - ax
        dy = cy[None, :] - ay
        dz = cz[None, :] - az
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        ok = r > 0
        rs = np.where(ok, r, 1)
        dotv = dx / rs
        Q = k * np.sqrt(np.maximum(0, 2 * (1 - dotv)))
        ss = (s_fac * Q) ** 2
        f0 = c_[a0:a1, None] + a_[a0:a1, 0][:, None] * np.exp(-b_[a0:a1, 0][:, None] * ss)
        for i in range(1, 5):
            f0 = f0 + a_[a0:a1, i][:, None] * np.exp(-b_[a0:a1, i][:, None] * ss)
        s_re = f0 + sanr[a0:a1, None]
        s_im = np.broadcast_to(sani[a0:a1, None], s_re.shape)
        if remove_forward:
            cut = Q < qc[None, :]
            s_re = np.where(cut, s_re - (f0z[a0:a1, None] + sanr[a0:a1, None]), s_re)
            s_im = np.where(cut, s_im - sani[a0:a1, None], s_im)
        ar = amp_r[a0:a1, None]
        ai = amp_i[a0:a1, None]
        t_re = ar * s_re - ai * s_im
        t_im = ar * s_im + ai * s_re
        ph = np.fmod(k * (ax + r), two_pi)
        c, s = np.cos(ph), np.sin(ph)
        v_re = t_re * c - t_im * s
        v_im = t_re * s + t_im * c
        scale = np.full(r.shape, rE, dtype=dt)
        if apply_decay:
            scale = scale * np.where(R0[None, :] > 0, R0[None, :] / rs, 1)
        if apply_pol:
            er = (ex * dx + ey * dy + ez * dz) / rs
            P = np.clip(dt(0.5) * (1 - pdeg) * (1 + dotv * dotv) + pdeg * (1 - er * er), 0, 1)
            scale = scale * np.sqrt(P)
        if analyser is not None and analyser[0]:
            du = np.stack([dx, dy, dz], axis=-1) / rs[..., None]
            scale = scale * analyser_weight(analyser[0], du.astype(np.float64), analyser[1],
                                            analyser[2], analyser[3]).astype(dt)
        acc_r += np.where(ok, v_re * scale, 0).sum(axis=0)
        acc_i += np.where(ok, v_im * scale, 0).sum(axis=0)
    return acc_r + 1j * acc_i


_LIBS = {}


def _sig(fn):
    d, f, i = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    fn.restype = i
    fn.argtypes = [i, f, f, f, i, f, f, f, f, i, i, f, f, f, d, i, d, d, d, d, i, f, f]
    return fn


def _load(path):
    if path not in _LIBS:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
        _LIBS[path] = ctypes.CDLL(path)
    return _LIBS[path]


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def kin_sum_c(pos_m, f0p, f0z, sanr, sani, amp_r, amp_i, cx, cy, cz, k_val,
              remove_forward, Ny, Nz, apply_pol, pol_deg, pol_e, apply_decay,
              precision="f64", threads=0):
    """C restatement (kin_sum.c); ``precision`` is "f64" or "f80"."""
    lib = _load(LIB_PATH)
    fn = getattr(lib, f"oracle_kin_sum_{precision}")
    fn.restype = ctypes.c_int
    d, f, i = ctypes.c_double, ctypes.c_void_p, ctypes.c_int
    fn.argtypes = [i, f, f, f, i, f, f, f, f, i, i, f, f, f, d, i, d, d, d, d, i, f, f, i]
    arrs = [_c(pos_m, np.float64), _c(f0p, np.float32), _c(f0z, np.float32),
            _c(sanr, np.float32), _c(sani, np.float32), _c(amp_r, np.float32),
            _c(amp_i, np.float32), _c(cx, np.float64), _c(cy, np.float64),
            _c(cz, np.float64)]
    P = Ny * Nz
    out_r = np.zeros(P, dtype=np.float64)
    out_i = np.zeros(P, dtype=np.float64)
    st = fn(int(arrs[0].shape[0]), _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
            int(bool(remove_forward)), _ptr(arrs[3]), _ptr(arrs[4]), _ptr(arrs[5]),
            _ptr(arrs[6]), int(Ny), int(Nz), _ptr(arrs[7]), _ptr(arrs[8]), _ptr(arrs[9]),
            float(k_val), int(bool(apply_pol)), float(pol_deg), float(pol_e[0]),
            float(pol_e[1]), float(pol_e[2]), int(bool(apply_decay)), _ptr(out_r),
            _ptr(out_i), int(threads))
    if st != 0:
        raise RuntimeError(f"oracle_kin_sum_{precision} returned {st}")
    return out_r + 1j * out_i


def ref_available():
    return os.path.exists(REF_LIB_PATH)


def ref_kin_sum(pos_m, f0p, f0z, sanr, sani, amp_r, amp_i, cx, cy, cz, k_val,
                remove_forward, Ny, Nz, apply_pol, pol_deg, pol_e, apply_decay,
                threads=1, slice_atoms=None):
    """The reference's compute_scattering_cffi, compiled from Beam.py's text.

    Driven as interact_beam_cpu drives it (Beam.py:5873-5955): atom slices of
    at least 1000 atoms run on a thread pool (ctypes releases the GIL), each
    slice writes float32 outputs, and the slices are summed in slice order.
    Returns complex128 sums of the float32 slice outputs.
    """
    from concurrent.futures import ThreadPoolExecutor
    lib = _load(REF_LIB_PATH)
    fn = _sig(lib.compute_scattering_cffi)
    pos = _c(pos_m, np.float64)
    f0p = _c(f0p, np.float32)
    tabs = [_c(x, np.float32) for x in (f0z, sanr, sani, amp_r, amp_i)]
    cxs = [_c(x, np.float64) for x in (cx, cy, cz)]
    n = pos.shape[0]
    P = Ny * Nz
    if slice_atoms is None:
        slice_atoms = max(1000, -(-n // max(1, threads)))

    def work(s0, s1):
        o_r = np.zeros(P, dtype=np.float32)
        o_i = np.zeros(P, dtype=np.float32)
        if s1 <= s0:
            return o_r, o_i
        sl = [np.ascontiguousarray(t[s0:s1]) for t in tabs]
        pp = np.ascontiguousarray(pos[s0:s1])
        ff = np.ascontiguousarray(f0p[s0:s1])
        st = fn(int(s1 - s0), _ptr(pp), _ptr(ff), _ptr(sl[0]), int(bool(remove_forward)),
                _ptr(sl[1]), _ptr(sl[2]), _ptr(sl[3]), _ptr(sl[4]), int(Ny), int(Nz),
                _ptr(cxs[0]), _ptr(cxs[1]), _ptr(cxs[2]), float(k_val),
                int(bool(apply_pol)), float(pol_deg), float(pol_e[0]), float(pol_e[1]),
                float(pol_e[2]), int(bool(apply_decay)), _ptr(o_r), _ptr(o_i))
        if st != 0:
            raise MemoryError("compute_scattering_cffi could not allocate")
        return o_r, o_i

    bounds = [(s, min(s + slice_atoms, n)) for s in range(0, n, slice_atoms)]
    total = np.zeros(P, dtype=np.complex64)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        for o_r, o_i in ex.map(lambda b: work(*b), bounds):
            total += (o_r + 1j * o_i).astype(np.complex64)
    return total.astype(np.complex128)


def ein_bilinear_numpy(pos_lab_f32, E0, du, dv, uc, vc):
    """Entrance amplitude, E0-only case of _ein_bilinear_cpu (Beam.py:956-1045).

    With tau = phi = 0 the reference's Ein reduces to the bilinear sample of
    E0 at the atom's beam coordinates (u = lab y, v = lab z for a +x beam),
    zero outside the grid.  Same float32 operations in the same order.
    """
    pos = np.asarray(pos_lab_f32, dtype=np.float32)
    N = pos.shape[0]
    out = np.zeros(N, dtype=np.complex64)
    NyB, NzB = E0.shape
    au = pos[:, 1]
    av = pos[:, 2]
    iu = au / float(du) + float(uc)
    iv = av / float(dv) + float(vc)
    inb = (iu >= 0.0) & (iu <= (NyB - 1)) & (iv >= 0.0) & (iv <= (NzB - 1))
    if not np.any(inb):
        return out
    iu_in, iv_in = iu[inb], iv[inb]
    i0 = np.floor(iu_in).astype(np.int64)
    j0 = np.floor(iv_in).astype(np.int64)
    i1 = np.clip(i0 + 1, 0, NyB - 1)
    j1 = np.clip(j0 + 1, 0, NzB - 1)
    fu = (iu_in - i0).astype(np.float32)
    fv = (iv_in - j0).astype(np.float32)
    w00 = (1.0 - fu) * (1.0 - fv)
    w01 = (1.0 - fu) * fv
    w10 = fu * (1.0 - fv)
    w11 = fu * fv
    E = np.asarray(E0, dtype=np.complex64).ravel()
    s = (E[i0 * NzB + j0] * w00 + E[i0 * NzB + j1] * w01 + E[i1 * NzB + j0] * w10
         + E[i1 * NzB + j1] * w11)
    out[inb] = s.astype(np.complex64)
    return out
