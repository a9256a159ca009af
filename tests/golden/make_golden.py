"""Write golden vectors of the kinematic sum from the reference's own C kernel.

Runs ``compute_scattering_cffi`` as compiled by ``oracle/build_ref.py`` from
the reference's source text (oracle/_ref/libref_scatter.so) on small seeded
inputs and stores inputs and float32 outputs in ``kin_golden.npz``.  Only the
numbers are committed.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mosaic_amd import tables  # noqa: E402
from mosaic_amd.detector import detector  # noqa: E402
from oracle import kinematic as O  # noqa: E402

K_10KEV = 2.0 * np.pi / tables.wavelength_from_energy_ev(10000.0)


def _species(labels, codes):
    c = tables.species_constants(labels, 10000.0)
    return c["wk"][codes], c["f0z"][codes], c["fp"][codes], c["fpp"][codes]


def _pixels(n_y, n_z, pitch_a, dist_a, two_theta_deg, eta_deg=0.0):
    d = detector()
    d.create_detector((n_y, n_z), (pitch_a, pitch_a))
    d.position_detector_absolute(dist_a, two_theta_deg, eta_deg)
    return np.asarray(d.pixel_coordinates, dtype=np.float64) / 1e10


def cases():
    rng = np.random.default_rng(20240611)
    out = []
    # A: three atoms of three species, complex amplitudes, 4 x 3 pixels at 2 cm
    pos = np.array([[0.0, 0.0, 0.0], [1.3, -0.7, 2.1], [-2.2, 0.4, -1.0]]) * 1e-10
    codes = np.array([0, 1, 2])
    pc = _pixels(4, 3, 2.0e5, 2.0e8, 25.0, 3.0)
    amp = np.array([1.0 + 0.0j, 0.6 - 0.3j, 0.2 + 0.9j], dtype=np.complex64)
    out.append(("three_atoms", pos, ["Au", "Sr", "O"], codes, amp, pc, 4, 3))
    # B: 64 random atoms within 20 A, 8 x 8 pixels at 1 m near 2theta = 30 deg
    pos = (rng.random((64, 3)) - 0.5) * 40e-10
    codes = rng.integers(0, 2, 64)
    pc = _pixels(8, 8, 7.5e5, 1.0e10, 30.0)
    amp = np.ones(64, dtype=np.complex64)
    out.append(("random64", pos, ["Au", "Ti"], codes, amp, pc, 8, 8))
    # C: forward-scattering corner: 9 atoms, 5 x 5 pixels around 2theta = 0
    pos = (rng.random((9, 3)) - 0.5) * 10e-10
    codes = np.zeros(9, dtype=np.int64)
    pc = _pixels(5, 5, 1.0e5, 1.0e9, 0.0)
    amp = (rng.random(9) + 1j * rng.random(9)).astype(np.complex64)
    out.append(("forward25", pos, ["Si"], codes, amp, pc, 5, 5))
    # D: near field, 6 atoms spread over +-6 A, 5 x 4 pixels 30 nm away, so that
    # the 1/r decay and the per-atom scattering angle matter
    pos = (rng.random((6, 3)) - 0.5) * 12e-10
    codes = np.array([0, 1, 0, 1, 0, 1])
    pc = _pixels(5, 4, 40.0, 300.0, 40.0, 10.0)
    amp = np.ones(6, dtype=np.complex64)
    out.append(("near30nm", pos, ["Au", "Ti"], codes, amp, pc, 5, 4))
    return out


FLAGS = [(False, False, False), (True, False, False), (False, True, False),
         (False, False, True), (True, True, True)]
POL = [(1.0, 0.0, 1.0, 0.0), (0.3, 0.0, 0.0, 1.0)]


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref_scatter.so is missing: run oracle/build_ref.py")
    data = {}
    for name, pos, labels, codes, amp, pc, ny, nz in cases():
        f0p, f0z, fp, fpp = _species(labels, codes)
        data[f"{name}/pos"] = pos
        data[f"{name}/f0p"] = f0p
        data[f"{name}/f0z"] = f0z
        data[f"{name}/sanr"] = fp
        data[f"{name}/sani"] = fpp
        data[f"{name}/amp"] = amp
        data[f"{name}/coords"] = pc
        data[f"{name}/shape"] = np.array([ny, nz])
        for fi, (rf, pol, dec) in enumerate(FLAGS):
            for pi, pv in enumerate(POL):
                res = O.ref_kin_sum(pos, f0p, f0z, fp, fpp, amp.real, amp.imag, pc[0], pc[1],
                                    pc[2], K_10KEV, rf, ny, nz, pol, pv[0], pv[1:], dec,
                                    threads=1)
                data[f"{name}/out_f{fi}_p{pi}"] = res.astype(np.complex64)
    data["k"] = np.array(K_10KEV)
    data["flags"] = np.array(FLAGS)
    data["pol"] = np.array(POL)
    np.savez_compressed(os.path.join(HERE, "kin_golden.npz"), **data)
    print("wrote", os.path.join(HERE, "kin_golden.npz"), len(data), "arrays")


if __name__ == "__main__":
    main()
