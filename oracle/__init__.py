"""CPU oracle for the kinematic scattering sum and the propagation path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``mosaic_amd`` imports this package:
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg use it, as the checker.

Contents:
  kinematic.py   -- NumPy restatement of compute_scattering_cffi
                    (Beam.py:663-816) in float64 or long double, plus the
                    ctypes binding of the C restatement (kin_sum.c) and of
                    the reference's own C kernel built into oracle/_ref/.
  propagation.py -- NumPy restatement of the angular-spectrum propagator,
                    the pad rule and the thin lens (Beam.py:4622-4900,
                    7945-8022; Optics.py:49-130, 513-639).
  adjoint.py     -- the displacement gradient from long-double atom terms
                    and its central finite differences.
  defects.py     -- NumPy twin of the device defect field.
The closed forms that pin these restatements live in the tests
(tests/test_oracle.py, tests/test_oracle_pins.py,
tests/test_propagation_oracle.py, tests/test_adjoint_oracle.py).
"""
