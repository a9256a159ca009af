"""Oracle for the optics elements and the detector PSF -- TEST
INFRASTRUCTURE ONLY.

NumPy restatements, in the reference's float32 expressions and order, of the
two-bounce Bragg magnifier (Optics.py:132-209), the k-space angular filter
(Optics.py:211-336), the Fresnel zone plate (Optics.py:338-425), the
aperture (Optics.py:427-471) and the detector PSF convolution
(Detector.py:873-1087, scipy.ndimage.convolve in float64).
"""
import numpy as np


def bragg_magnifier(F, Mx, My, reflectivity=1.0, phase_shift=0.0, order=1, pad_mode="zeros",
                    conserve_energy=True):
    F = np.asarray(F, np.complex64)
    Ny, Nx = F.shape
    cx, cy = (Nx - 1) / 2.0, (Ny - 1) / 2.0
    Xi = cx + (np.arange(Nx, dtype=np.float32)[None, :] - cx) / Mx
    Yi = cy + (np.arange(Ny, dtype=np.float32)[:, None] - cy) / My
    if order == 0:
        xn, yn = np.rint(Xi).astype(np.int32), np.rint(Yi).astype(np.int32)
        valid = (xn >= 0) & (xn < Nx) & (yn >= 0) & (yn < Ny)
        out = F[np.clip(yn, 0, Ny - 1), np.clip(xn, 0, Nx - 1)]
        out = np.where(valid, out, np.complex64(0.0))
    else:
        x0, y0 = np.floor(Xi).astype(np.int32), np.floor(Yi).astype(np.int32)
        wx, wy = (Xi - x0).astype(np.float32), (Yi - y0).astype(np.float32)
        if pad_mode == "edge":
            x0c, x1c = np.clip(x0, 0, Nx - 1), np.clip(x0 + 1, 0, Nx - 1)
            y0c, y1c = np.clip(y0, 0, Ny - 1), np.clip(y0 + 1, 0, Ny - 1)
            wx, wy = Xi - x0c, Yi - y0c
            valid = None
        else:
            valid = (Xi >= 0) & (Xi < Nx - 1) & (Yi >= 0) & (Yi < Ny - 1)
            x0c = np.clip(x0, 0, Nx - 2)
            x1c = x0c + 1
            y0c = np.clip(y0, 0, Ny - 2)
            y1c = y0c + 1
        out = (F[y0c, x0c] * (1 - wx) * (1 - wy) + F[y0c, x1c] * wx * (1 - wy)
               + F[y1c, x0c] * (1 - wx) * wy + F[y1c, x1c] * wx * wy).astype(np.complex64)
        if valid is not None:
            out = np.where(valid, out, np.complex64(0.0))
    amp = np.sqrt(max(reflectivity, 0.0))
    if conserve_energy and abs(Mx * My) > 0:
        amp *= 1.0 / np.sqrt(abs(Mx * My))
    return (out * np.complex64(amp * np.exp(1j * phase_shift))).astype(np.complex64)


def angular_mask(Ny, Nx, dx, dy, wavelength, filt):
    k0 = 2.0 * np.pi / float(wavelength)
    kx = (2.0 * np.pi) * np.fft.fftfreq(Nx, d=float(dx)).astype(np.float32)
    ky = (2.0 * np.pi) * np.fft.fftfreq(Ny, d=float(dy)).astype(np.float32)
    tx, ty = (kx / k0)[None, :], (ky / k0)[:, None]
    cx = float(filt.get("center_x_mrad", 0.0)) * 1e-3
    cy = float(filt.get("center_y_mrad", 0.0)) * 1e-3
    hx = float(filt.get("half_angle_x_mrad", 0.0)) * 1e-3
    hy = float(filt.get("half_angle_y_mrad", hx * 1e3)) * 1e-3
    mode = str(filt.get("mode", "2d")).lower()
    shape = str(filt.get("shape", "circular")).lower()
    rolloff = str(filt.get("rolloff", "tophat")).lower()
    phi = float(np.deg2rad(float(filt.get("roll_deg", 0.0))))
    tX, tY = tx - cx, ty - cy
    c, s = float(np.cos(phi)), float(np.sin(phi))
    tpar, tperp = c * tX + s * tY, -s * tX + c * tY
    eps = 1e-30
    order = max(1, int(filt.get("order", 4)))
    if mode == "1d":
        sv = tpar / max(hx, eps)
        H = (1.0 / (1.0 + sv * sv) if rolloff == "darwin" else
             (np.abs(sv) <= 1.0).astype(np.float32) if rolloff == "tophat" else
             1.0 / np.sqrt(1.0 + np.power(np.abs(sv), 2 * order)))
    elif shape == "elliptical":
        r2 = (tpar / max(hx, eps)) ** 2 + (tperp / max(hy, eps)) ** 2
        H = (1.0 / (1.0 + r2) if rolloff == "darwin" else
             (r2 <= 1.0).astype(np.float32) if rolloff == "tophat" else
             1.0 / np.sqrt(1.0 + np.power(r2, order)))
    else:
        delta = np.sqrt(tX * tX + tY * tY)
        hw = max(hx, eps)
        H = (1.0 / (1.0 + (delta / hw) * (delta / hw)) if rolloff == "darwin" else
             (delta <= hw).astype(np.float32) if rolloff == "tophat" else
             1.0 / np.sqrt(1.0 + np.power(delta / hw, 2 * order)))
    return np.broadcast_to(H, (Ny, Nx)).astype(np.float32)


def angular_filter(E, dx, dy, wavelength, filt):
    E = np.asarray(E, np.complex64)
    H = angular_mask(E.shape[0], E.shape[1], dx, dy, wavelength, filt)
    amp = float(np.sqrt(max(float(filt.get("transmission", 1.0)), 0.0)))
    ph = np.complex64(amp * np.exp(1j * float(filt.get("phase_shift", 0.0))))
    return (np.fft.ifft2(np.fft.fft2(E) * H) * ph).astype(np.complex64)


def zone_plate(E, dx, dy, wavelength, zp):
    E = np.asarray(E, np.complex64)
    Ny, Nx = E.shape
    drN, D = float(zp["outermost_zone_width_nm"]) * 1e-9, float(zp["diameter_um"]) * 1e-6
    order = max(1, int(zp.get("order", 1)))
    f = D * drN / wavelength
    f_work, R = f / order, D / 2.0
    amp = float(np.sqrt(max(float(zp.get("efficiency", 1.0)), 0.0)))
    cs = float(zp.get("central_stop_fraction", 0.0))
    typ = str(zp.get("zone_plate_type", "amplitude")).lower()
    x = (np.arange(Nx, dtype=np.float32) - (Nx - 1) / 2.0) * float(dx)
    y = (np.arange(Ny, dtype=np.float32) - (Ny - 1) / 2.0) * float(dy)
    R2 = x[None, :] ** 2 + y[:, None] ** 2
    parity = np.floor(R2 / (float(wavelength) * f)).astype(np.int32) % 2
    inside = R2 <= R * R
    if cs > 0.0:
        inside &= R2 >= (R * cs) ** 2
    T = np.zeros((Ny, Nx), np.complex64)
    if typ == "ideal":
        phase = (-0.5 * ((2.0 * np.pi / wavelength) / f_work)) * R2
        T[inside] = amp * np.exp(1j * phase[inside]).astype(np.complex64)
    elif typ == "phase":
        T[inside & (parity == 0)] = amp
        T[inside & (parity == 1)] = -amp
    else:
        T[inside & (parity == 0)] = amp
    taper = float(zp.get("edge_taper", 0.0))
    if taper > 0.0:
        r = np.sqrt(R2)
        r_in = R * (1.0 - taper)
        tz = (r > r_in) & (r <= R)
        T[tz] *= (0.5 * (1.0 + np.cos(np.pi * (r[tz] - r_in) / (R * taper)))).astype(np.float32)
    return (E * T).astype(np.complex64)


def aperture(E, dx, dy, width_mm, shape="square"):
    E = np.array(E, np.complex64, copy=True)
    Ny, Nx = E.shape
    half = 0.5 * width_mm * 1e-3
    x = (np.arange(Nx, dtype=np.float32) - (Nx - 1) / 2.0) * float(dx)
    y = (np.arange(Ny, dtype=np.float32) - (Ny - 1) / 2.0) * float(dy)
    if shape == "circular":
        mask = (x[None, :] ** 2 + y[:, None] ** 2) <= half * half
    else:
        mask = (np.abs(x) <= half)[None, :] & (np.abs(y) <= half)[:, None]
    E[~mask] = 0.0
    return E


def psf(E2d, kernel, mode="constant", field="intensity"):
    from scipy.ndimage import convolve
    E2d = np.asarray(E2d, np.complex64)
    data = (np.abs(E2d) ** 2 if field == "intensity" else np.abs(E2d)).astype(np.float64)
    conv = np.maximum(convolve(data, kernel, mode=mode), 0.0)
    amp = (np.sqrt(conv) if field == "intensity" else conv).astype(np.float32)
    ph = np.angle(E2d)
    return (amp * (np.cos(ph) + 1j * np.sin(ph))).astype(np.complex64)
