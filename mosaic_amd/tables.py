"""Scattering-factor tables and the per-species constants of the kinematic sum.

The numbers come from ``data/scattering_tables.npz`` (built by
``data/make_tables.py`` from the DABAX Waasmaier-Kirfel and Cromer-Liberman
files).  The lookups follow the reference:

* ``parse_f0_db_all`` -> dict symbol -> (11,) float32 ``[a1..a5, c, b1..b5]``
  (Beam.py:503-527);
* ``parse_f1f2_db_all`` -> dict symbol -> (n, 3) float32 ``[E_eV, f1, f2]``
  (Beam.py:529-568);
* ``get_f1f2_from_params`` -> piecewise-linear f' + i f'' in energy with the
  end segments extended (Beam.py:570-598);
* ``build_f0_zero_dict`` -> f0(0) = c + sum(a_i) (Beam.py:600-615).
"""
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_TABLE_PATH = os.path.join(_HERE, "data", "scattering_tables.npz")
_CACHE = {}

# SI constants used by beam.create_beam (Beam.py:90-94).
H_PLANCK = 6.62607015e-34
C_LIGHT = 299792458
Q_E = 1.602176634e-19
R_E = 2.81794092e-15          # classical electron radius [m] (Beam.py:703)


def wavelength_from_energy_ev(energy_ev):
    """Wavelength [m] for a photon energy in eV, as beam.create_beam computes it."""
    return (H_PLANCK / Q_E) * C_LIGHT / float(energy_ev)


def _load():
    if "npz" not in _CACHE:
        with np.load(_TABLE_PATH, allow_pickle=False) as z:
            _CACHE["npz"] = {k: z[k] for k in z.files}
    return _CACHE["npz"]


def parse_f0_db_all():
    """Element symbol -> float32 (11,) Waasmaier-Kirfel row."""
    if "f0" not in _CACHE:
        z = _load()
        _CACHE["f0"] = {str(s): z["wk_params"][i].copy()
                        for i, s in enumerate(z["wk_symbols"])}
    return _CACHE["f0"]


def parse_f1f2_db_all():
    """Element symbol -> float32 (n, 3) Cromer-Liberman table [E_eV, f1, f2]."""
    if "f1f2" not in _CACHE:
        z = _load()
        off = z["cl_offsets"]
        _CACHE["f1f2"] = {str(s): z["cl_rows"][off[i]:off[i + 1]].copy()
                          for i, s in enumerate(z["cl_symbols"])}
    return _CACHE["f1f2"]


def get_f1f2_from_params(energy, f1f2_table):
    """f1 + i f2 at ``energy`` [eV], linear in E, end segments extended.

    Same index rule and arithmetic as Beam.py:570-598 (float32 table values,
    Python-float energy).
    """
    E = energy
    energies = f1f2_table[:, 0]
    idx = int(np.searchsorted(energies, E))
    if idx >= len(energies):
        idx = len(energies) - 1
    if idx == 0:
        idx = 1
    E0, f10, f20 = energies[idx - 1], f1f2_table[idx - 1, 1], f1f2_table[idx - 1, 2]
    E1, f11, f21 = energies[idx], f1f2_table[idx, 1], f1f2_table[idx, 2]
    denom = (E1 - E0) if (E1 > E0) else 1e-20
    w = (E - E0) / denom
    f1 = f10 + (f11 - f10) * w
    f2 = f20 + (f21 - f20) * w
    return complex(f1 + 1j * f2)


def build_f0_zero_dict(db_f0):
    """Element -> f0(0) = c + a1 + ... + a5."""
    out = {}
    for el, p in db_f0.items():
        out[el] = float(p[5] + p[0] + p[1] + p[2] + p[3] + p[4])
    return out


def species_constants(labels, energy_ev):
    """Per-species arrays for a label table.

    Unknown symbols get all-zero constants, which makes their atoms scatter
    nothing -- the reference skips them the same way (Beam.py:5689-5692).

    Returns:
        dict with ``wk`` (S, 11) float32, ``f0z`` (S,) float32, ``fp`` (S,)
        float32 (f'), ``fpp`` (S,) float32 (f'').
    """
    db_f0 = parse_f0_db_all()
    db_f12 = parse_f1f2_db_all()
    f0z_d = build_f0_zero_dict(db_f0)
    S = len(labels)
    wk = np.zeros((S, 11), dtype=np.float32)
    f0z = np.zeros(S, dtype=np.float32)
    fp = np.zeros(S, dtype=np.float32)
    fpp = np.zeros(S, dtype=np.float32)
    for i, el in enumerate(labels):
        el = str(el)
        if el not in db_f0:
            continue
        wk[i] = db_f0[el]
        f0z[i] = np.float32(f0z_d.get(el, 0.0))
        tbl = db_f12.get(el)
        if tbl is not None:
            c = get_f1f2_from_params(energy_ev, tbl)
            fp[i] = np.float32(c.real)
            fpp[i] = np.float32(c.imag)
    return {"wk": wk, "f0z": f0z, "fp": fp, "fpp": fpp}


def polarization_from_rate(rho):
    """(p, ex, ey, ez) for a +x beam from the perpendicular fraction rho.

    rho >= 0.5 -> p = 2 rho - 1 along lab y; else p = 1 - 2 rho along lab z
    (Beam.py:31-47).
    """
    rho = float(np.clip(rho, 0.0, 1.0))
    p = abs(2.0 * rho - 1.0)
    return (p, 0.0, 1.0, 0.0) if rho >= 0.5 else (p, 0.0, 0.0, 1.0)
