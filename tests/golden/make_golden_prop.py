"""Golden vectors of the propagation multiplier from the reference's own
prop_mul_cpu (Beam.py:4853-4891, built into oracle/_ref by build_ref.py).

    python tests/golden/make_golden_prop.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import propagation as P  # noqa: E402


def main():
    if not os.path.exists(P.REF_PROP):
        raise SystemExit("oracle/_ref/libref_prop.so is missing: run oracle/build_ref.py")
    rng = np.random.default_rng(7)
    lam = 1.2398419843320025e-10
    k = np.float32(2 * np.pi / lam)
    data = {"k": np.array(k)}
    # (Ny, Nz, d [m], z [m]): far inside the propagating band, and a grid finer
    # than lambda / 2 whose outer frequencies are evanescent
    cases = [(48, 40, 2.0e-8, 3.0e-4), (32, 36, 4.0e-11, 2.0e-10), (40, 32, 4.0e-11, -3.0e-10)]
    for i, (ny, nz, d, z) in enumerate(cases):
        ky = ((2.0 * np.pi) * np.fft.fftfreq(ny, d=d).astype(np.float32)).astype(np.float32)
        kz = ((2.0 * np.pi) * np.fft.fftfreq(nz, d=d).astype(np.float32)).astype(np.float32)
        F = (rng.standard_normal((nz, ny)) + 1j * rng.standard_normal((nz, ny))).astype(np.complex64)
        data[f"c{i}/ky"] = ky
        data[f"c{i}/kz"] = kz
        data[f"c{i}/z"] = np.array(np.float32(z))
        data[f"c{i}/F"] = F
        data[f"c{i}/out"] = P.ref_prop_mul(F, ky, kz, k, np.float32(z))
    np.savez_compressed(os.path.join(HERE, "prop_golden.npz"), **data)
    print("wrote prop_golden.npz", len(cases), "cases")


if __name__ == "__main__":
    main()
