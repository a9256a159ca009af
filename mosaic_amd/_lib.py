"""ctypes binding of ``libmosaic_amd.so`` (the C ABI in include/mosaic_amd.h).

The library must be built (``python -m mosaic_amd.build``); a missing
library raises ``MosaicLibraryError`` -- there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmosaic_amd.so")

MOSAIC_OK = 0
MOSAIC_ERR_ARGS = 1
MOSAIC_ERR_HIP = 2
MOSAIC_ERR_UNSUPPORTED = 3
MOSAIC_ERR_WORKSPACE = 4
MOSAIC_ERR_TOO_LARGE = 5

_ERRORS = {
    MOSAIC_ERR_ARGS: "invalid argument",
    MOSAIC_ERR_HIP: "HIP runtime error",
    MOSAIC_ERR_UNSUPPORTED: "declined: geometry outside the fast path's error bound",
    MOSAIC_ERR_WORKSPACE: "workspace too small",
    MOSAIC_ERR_TOO_LARGE: "beyond the allocation-free call's fixed scratch",
}

# Every symbol include/mosaic_amd.h declares.
EXPORTS = (
    "mosaic_kin_workspace",
    "mosaic_kin_stage",
    "mosaic_kin_restage",
    "mosaic_kin_pixels",
    "mosaic_kin_sum",
    "mosaic_kin_stats",
    "mosaic_kin_timing",
    "mosaic_kin_adjoint",
    "mosaic_kin_adjoint_scratch",
    "mosaic_scatter_workspace",
    "mosaic_scatter",
    "mosaic_scatter_noalloc",
    "mosaic_sincos_rev",
    "mosaic_mbeam_amplitude",
    "mosaic_ein_bilinear",
    "mosaic_ein_depth_workspace",
    "mosaic_ein_depth_bounds",
    "mosaic_ein_column_map",
    "mosaic_ein_depth",
    "mosaic_asp_workspace",
    "mosaic_asp_propagate",
    "mosaic_lens_apply",
    "mosaic_bragg_magnifier",
    "mosaic_angular_filter_workspace",
    "mosaic_angular_filter",
    "mosaic_zone_plate",
    "mosaic_aperture",
    "mosaic_psf_workspace",
    "mosaic_psf_apply",
    "mosaic_defect_positions",
    "mosaic_sg_gemm",
    "mosaic_sg_bias_act",
    "mosaic_sg_act_bwd",
    "mosaic_sg_bias_grad",
    "mosaic_sg_shuffle2",
    "mosaic_sg_mse",
    "mosaic_sg_adam",
    "mosaic_sg_ensemble_stats",
    "mosaic_sg_pack",
    "mosaic_sg_infer_workspace",
    "mosaic_sg_infer",
)


class MosaicLibraryError(RuntimeError):
    pass


class MosaicDeclined(RuntimeError):
    """The fast path declined (MOSAIC_ERR_UNSUPPORTED); nothing was written."""


class Species(ctypes.Structure):
    _fields_ = [("wk", ctypes.c_float * 11), ("f0_zero", ctypes.c_float),
                ("fp", ctypes.c_float), ("fpp", ctypes.c_float)]


class KinOpts(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("remove_forward", ctypes.c_int),
                ("apply_pol", ctypes.c_int), ("apply_decay", ctypes.c_int),
                ("path", ctypes.c_int), ("pol_deg", ctypes.c_double),
                ("pol_e", ctypes.c_double * 3), ("analyser_kind", ctypes.c_int),
                ("gm_pixel_eval", ctypes.c_int), ("analyser_axis", ctypes.c_double * 3),
                ("analyser_accept_rad", ctypes.c_double),
                ("analyser_darwin_hw_rad", ctypes.c_double)]


class Lens(ctypes.Structure):
    _fields_ = [("focal_length_m", ctypes.c_double), ("parab_coef", ctypes.c_double),
                ("uniform_amp", ctypes.c_float), ("pad_", ctypes.c_float)]


class AngularFilter(ctypes.Structure):
    _fields_ = [("center_x_mrad", ctypes.c_double), ("center_y_mrad", ctypes.c_double),
                ("half_angle_x_mrad", ctypes.c_double), ("half_angle_y_mrad", ctypes.c_double),
                ("transmission", ctypes.c_double), ("phase_shift", ctypes.c_double),
                ("roll_deg", ctypes.c_double), ("mode_1d", ctypes.c_int),
                ("elliptical", ctypes.c_int), ("rolloff", ctypes.c_int), ("order", ctypes.c_int)]


class SgNet(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("W1", "b1", "W2", "b2", "W3", "b3", "D1", "d1",
                                                "D2", "d2", "D3", "d3", "D4", "d4")]


class ZonePlate(ctypes.Structure):
    _fields_ = [("outermost_zone_width_nm", ctypes.c_double), ("diameter_um", ctypes.c_double),
                ("efficiency", ctypes.c_double), ("central_stop_fraction", ctypes.c_double),
                ("edge_taper", ctypes.c_double), ("order", ctypes.c_int), ("type", ctypes.c_int)]


_LIB = None

_vp, _i, _i64, _d, _sz = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                          ctypes.c_size_t)
_ll, _f = ctypes.c_longlong, ctypes.c_float
_SIGS = {
    "mosaic_kin_workspace": [_i64, _i, _vp, _i, _i, _i, ctypes.POINTER(_sz)],
    "mosaic_kin_stage": [_i64, _vp, _vp, _vp, _i, _vp, _d, _i, _i, _i, _vp, _sz, _vp, _vp],
    "mosaic_kin_restage": [_i64, _vp, _vp, _i, _vp, _d, _i, _i, _i, _vp, _sz, _vp, _vp],
    "mosaic_kin_pixels": [_i, _vp, _vp, _i, _i, _vp, _vp, _sz, _vp, _vp],
    "mosaic_kin_sum": [_i64, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _i, _vp, _vp, _i, _vp, _vp,
                       _vp, _sz, _vp, _vp],
    "mosaic_kin_stats": [_i64, _i, _vp, _i, _i, _i, _vp, _sz, _vp, _vp, _vp],
    "mosaic_kin_timing": [_i, _vp],
    "mosaic_kin_adjoint": [_i64, _i, _vp, _d, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp,
                           _sz, _vp, _sz, _i, _vp],
    "mosaic_kin_adjoint_scratch": [_i64, _i, _vp, _i, _i, _i, ctypes.POINTER(_sz)],
    "mosaic_scatter_workspace": [_i, _i, _i, ctypes.POINTER(_sz)],
    "mosaic_scatter": [_i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _d, _i,
                       _d, _d, _d, _d, _i, _vp, _vp, _vp, _sz, _vp],
    "mosaic_scatter_noalloc": [_i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp,
                               _d, _i, _d, _d, _d, _d, _i, _vp, _vp, _vp],
    "mosaic_ein_bilinear": [_i64, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp, _vp],
    "mosaic_sincos_rev": [_i64, _vp, _vp, _vp, _vp],
    "mosaic_mbeam_amplitude": [_i64, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    "mosaic_ein_depth_workspace": [_i, _i, ctypes.POINTER(_sz)],
    "mosaic_ein_depth_bounds": [_i64, _vp, _vp, _vp, _i, _i, _vp, _vp, _sz, _vp],
    "mosaic_ein_column_map": [_i64, _vp, _vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _i,
                              _vp, _vp, _sz, _vp],
    "mosaic_ein_depth": [_i64, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _d, _d, _vp, _vp,
                         _sz, _vp],
    "mosaic_asp_workspace": [_i, _i, ctypes.POINTER(_sz)],
    "mosaic_asp_propagate": [_vp, _i, _i, _d, _d, _d, _d, _i, _i, _i, ctypes.c_float,
                             ctypes.c_float, _vp, _i, _vp, _vp, _sz, _vp],
    "mosaic_lens_apply": [_vp, _i, _i, _d, _d, _d, _vp, _vp],
    "mosaic_bragg_magnifier": [_vp, _vp, _i, _i, _d, _d, _d, _d, _i, _i, _i, _vp],
    "mosaic_angular_filter_workspace": [_i, _i, ctypes.POINTER(_sz)],
    "mosaic_angular_filter": [_vp, _i, _i, _d, _d, _d, _vp, _vp, _sz, _vp],
    "mosaic_zone_plate": [_vp, _i, _i, _d, _d, _d, _vp, _vp],
    "mosaic_aperture": [_vp, _i, _i, _d, _d, _d, _i, _vp],
    "mosaic_psf_workspace": [_i, _i, _i, ctypes.POINTER(_sz)],
    "mosaic_psf_apply": [_vp, _i, _i, _vp, _i, _i, _i, _i, _vp, _sz, _vp],
    "mosaic_defect_positions": [_i64, _vp, _vp, _vp, _vp],
    "mosaic_sg_gemm": [_i, _i, _i, _i, _i, _vp, _i, _ll, _vp, _i, _ll, _vp, _i, _ll, _i, _f, _f,
                       _vp],
    "mosaic_sg_bias_act": [_vp, _vp, _vp, _ll, _i, _i, _ll, _ll, _i, _i, _vp],
    "mosaic_sg_act_bwd": [_vp, _vp, _vp, _ll, _i, _vp],
    "mosaic_sg_bias_grad": [_vp, _vp, _ll, _i, _i, _ll, _ll, _i, _vp],
    "mosaic_sg_shuffle2": [_vp, _vp, _i, _i, _i, _i, _ll, _i, _i, _vp],
    "mosaic_sg_mse": [_vp, _vp, _vp, _vp, _ll, _ll, _i, _vp, _ll, _vp],
    "mosaic_sg_adam": [_vp, _vp, _vp, _vp, _ll, _f, _f, _f, _f, _i, _f, _vp],
    "mosaic_sg_ensemble_stats": [_vp, _vp, _vp, _i, _i, _i, _ll, _vp],
    "mosaic_sg_pack": [_i, _i, _i, _vp, _ll, _vp, _vp],
    "mosaic_sg_infer_workspace": [_i, _i, ctypes.POINTER(_sz)],
    "mosaic_sg_infer": [_i, _i, _vp, ctypes.POINTER(SgNet), _vp, _vp, _sz, _vp, _vp],
}


def lib():
    """The loaded library; raises MosaicLibraryError when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MosaicLibraryError(
                f"{LIB_PATH} is missing: build it with `python -m mosaic_amd.build` "
                "(there is no non-HIP fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _LIB = L
    return _LIB


def check(status, what):
    if status == MOSAIC_OK:
        return
    if status == MOSAIC_ERR_UNSUPPORTED:
        raise MosaicDeclined(f"{what}: {_ERRORS[status]}")
    raise MosaicLibraryError(f"{what} failed: {_ERRORS.get(status, status)}")
