"""The reference-side ctypes binding of the device compute_scattering_cffi.

This is the stub a MOSAIC maintainer adds next to
``compile_compute_scattering_cffi`` (Beam.py:629-700): the same argument
list as the CPU kernel, every array a device pointer, plus a workspace and a
stream.  The tests run it exactly as written here, with full ``argtypes``,
so that the binding shown in the integration notes is known to work.
"""
import ctypes

_i, _d, _vp, _sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t

ARGTYPES = [
    _i,          # atom_count
    _vp,         # positions (N, 3) float64 lab metres
    _vp,         # f0_params (N, 11) float32 WK rows
    _vp,         # f0_zero (N,) float32
    _i,          # remove_forward
    _vp, _vp,    # s_anom_real, s_anom_imag (N,) float32
    _vp, _vp,    # initial_amp_r, initial_amp_i (N,) float32
    _i, _i,      # Ny, Nz
    _vp, _vp, _vp,   # coords_x, coords_y, coords_z (Ny*Nz,) float64 metres
    _d,          # k_val [rad/m]
    _i,          # apply_pol
    _d, _d, _d, _d,  # pol_deg, pol_ex, pol_ey, pol_ez
    _i,          # apply_spherical_decay
    _vp, _vp,    # out_r, out_i (Ny*Nz,) float32, accumulated
    _vp, _sz,    # workspace, workspace_bytes
    _vp,         # stream (hipStream_t)
]


def bind(lib_path):
    """Load the library and declare the two entry points' signatures."""
    lib = ctypes.CDLL(lib_path)
    lib.mosaic_scatter_workspace.argtypes = [_i, _i, _i, ctypes.POINTER(_sz)]
    lib.mosaic_scatter_workspace.restype = _i
    lib.mosaic_scatter.argtypes = ARGTYPES
    lib.mosaic_scatter.restype = _i
    return lib


def workspace_bytes(lib, atom_count, Ny, Nz):
    n = _sz(0)
    rc = lib.mosaic_scatter_workspace(atom_count, Ny, Nz, ctypes.byref(n))
    if rc != 0:
        raise RuntimeError(f"mosaic_scatter_workspace returned {rc}")
    return n.value


def compute_scattering_gpu(lib, atom_count, positions, f0_params, f0_zero, remove_forward,
                           s_anom_real, s_anom_imag, initial_amp_r, initial_amp_i, Ny, Nz,
                           coords_x, coords_y, coords_z, k_val, apply_pol, pol_deg, pol_ex,
                           pol_ey, pol_ez, apply_spherical_decay, out_r, out_i, workspace,
                           workspace_bytes, stream=0):
    """compute_scattering_cffi's call, with integer device addresses for the
    arrays; raises where the reference returns 1 (Beam.py:5768-5769)."""
    rc = lib.mosaic_scatter(atom_count, positions, f0_params, f0_zero, int(remove_forward),
                            s_anom_real, s_anom_imag, initial_amp_r, initial_amp_i, Ny, Nz,
                            coords_x, coords_y, coords_z, float(k_val), int(apply_pol),
                            float(pol_deg), float(pol_ex), float(pol_ey), float(pol_ez),
                            int(apply_spherical_decay), out_r, out_i, workspace,
                            workspace_bytes, stream)
    if rc != 0:
        raise MemoryError(f"mosaic_scatter returned {rc}")
    return rc
