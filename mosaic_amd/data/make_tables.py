"""Build ``scattering_tables.npz`` from the DABAX text tables.

The kinematic sum needs two published tables:

* Waasmaier-Kirfel f0 fits (``f0_WaasKirf.dat``): one row of 11 floats
  ``[a1..a5, c, b1..b5]`` per element/ion, parsed the way
  ``beam.parse_f0_db_all`` does (Beam.py:503-527): a ``#S`` line names the
  symbol (third whitespace field), the next non-comment line with exactly 11
  numbers is its row, a later row for the same symbol replaces it.
* Cromer-Liberman anomalous factors (``f1f2_CromerLiberman.dat``): rows of
  ``[E_eV, f1, f2]`` per element, parsed as ``beam.parse_f1f2_db_all`` does
  (Beam.py:529-568): rows of exactly three numbers between ``#S`` lines.

Both are stored as float32, as the reference keeps them.  Run once where the
DABAX files exist::

    python mosaic_amd/data/make_tables.py /root/reference/databases/scattering

The output is data only (numbers and element symbols).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_f0(path):
    rows = {}
    element = None
    with open(path) as fh:
        for line in fh:
            if line.startswith("#S"):
                element = line.split()[2].strip()
            elif (not line.startswith("#")) and element is not None:
                vals = [float(x) for x in line.split()]
                if len(vals) == 11:
                    rows[element] = np.asarray(vals, dtype=np.float32)
    return rows


def parse_f1f2(path):
    tables = {}
    element = None
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#S"):
                if element is not None and rows:
                    tables[element] = np.asarray(rows, dtype=np.float32)
                element = line.split()[2].strip()
                rows = []
            elif (not line.startswith("#")) and element is not None:
                vals = [float(x) for x in line.split()]
                if len(vals) == 3:
                    rows.append(vals)
    if element is not None and rows:
        tables[element] = np.asarray(rows, dtype=np.float32)
    return tables


def main(src_dir, out=os.path.join(HERE, "scattering_tables.npz")):
    f0 = parse_f0(os.path.join(src_dir, "f0_WaasKirf.dat"))
    f12 = parse_f1f2(os.path.join(src_dir, "f1f2_CromerLiberman.dat"))
    wk_sym = sorted(f0)
    cl_sym = sorted(f12)
    offsets = np.zeros(len(cl_sym) + 1, dtype=np.int64)
    for i, s in enumerate(cl_sym):
        offsets[i + 1] = offsets[i] + f12[s].shape[0]
    np.savez_compressed(
        out,
        wk_symbols=np.asarray(wk_sym),
        wk_params=np.stack([f0[s] for s in wk_sym]).astype(np.float32),
        cl_symbols=np.asarray(cl_sym),
        cl_offsets=offsets,
        cl_rows=np.concatenate([f12[s] for s in cl_sym]).astype(np.float32),
    )
    print(f"wrote {out}: {len(wk_sym)} f0 rows, {len(cl_sym)} f1f2 tables "
          f"({int(offsets[-1])} rows)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/databases/scattering")
