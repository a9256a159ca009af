"""Operator-level kinematic sum: torch device buffers in, HIP kernels out.

This is the seam the reference calls ``dispatch`` (Beam.py:9394-9496): the
caller owns every buffer, the output is accumulated into, and a decline
(``MosaicDeclined``) writes nothing.  Device memory and streams come from
PyTorch; all arithmetic runs in ``libmosaic_amd.so``.
"""
import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import KinOpts, Species, check


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(int(s.cuda_stream))


def _ptr(t):
    return ctypes.c_void_p(int(t.data_ptr())) if t is not None else ctypes.c_void_p(0)


def species_table(consts):
    """ctypes array of mosaic_species_t from tables.species_constants()."""
    S = consts["wk"].shape[0]
    arr = (Species * S)()
    for s in range(S):
        for m in range(11):
            arr[s].wk[m] = float(consts["wk"][s, m])
        arr[s].f0_zero = float(consts["f0z"][s])
        arr[s].fp = float(consts["fp"][s])
        arr[s].fpp = float(consts["fpp"][s])
    return arr


PATHS = {"auto": 0, "tile": 1, "rows": 2, "mma": 3, "poly": 4, "exact": 5, "probe": 6,
         "gemm": 7}

# analyser_mode names of the reference's GPU driver (Beam.py:6134-6143)
_ANALYSER = {"off": 0, "none": 0, "disabled": 0, "top_hat": 1, "tophat": 1, "top-hat": 1,
             "top": 1, "darwin": 2, "rolloff": 2, "roll-off": 2}


def analyser_kind(mode):
    """0 (off), 1 (top-hat) or 2 (Darwin) from the reference's analyser_mode
    argument: a name, an int, or None; unknown values mean off
    (Beam.py:6134-6143)."""
    if mode is None:
        return 0
    if isinstance(mode, str):
        return _ANALYSER.get(mode.strip().lower(), 0)
    k = int(mode)
    return k if k in (0, 1, 2) else 0


def kin_opts(k, remove_forward=False, apply_pol=False, pol=(0.0, 0.0, 1.0, 0.0),
             apply_decay=False, path="auto", analyser=0, analyser_axis=(1.0, 0.0, 0.0),
             accept_rad=0.0, darwin_hw_rad=0.0, gm_pixel_eval=False):
    """mosaic_kin_opts_t; ``path`` is "auto", "exact" (per pair), "tile"
    (per-tile factors), "rows" (row recurrence), "mma" (matrix-core patches),
    "poly" (moment expansion over the same patches), "gemm" (the moment
    expansion with its per-atom work on the matrix cores; "auto" prefers it
    from 64 patches up) or "probe" (a few scattered pixels, one lane per
    frame and pixel; "auto" takes it for at most 16384 pixel-poses); all
    but "auto" and "exact" decline when their bounds do not hold.  ``analyser`` is the kind (0 off, 1
    top-hat, 2 Darwin) with its axis (lab), acceptance half-angle and Darwin
    half-width; it takes the exact kernel.  ``gm_pixel_eval`` makes "gemm"
    evaluate every (unit, patch) pixel by pixel, its evaluation's fallback
    from the separable phase model."""
    o = KinOpts()
    o.k = float(k)
    o.remove_forward = int(bool(remove_forward))
    o.apply_pol = int(bool(apply_pol))
    o.apply_decay = int(bool(apply_decay))
    o.path = PATHS[path]
    o.pol_deg = float(pol[0])
    for d in range(3):
        o.pol_e[d] = float(pol[1 + d])
    o.analyser_kind = int(analyser)
    for d in range(3):
        o.analyser_axis[d] = float(analyser_axis[d])
    o.analyser_accept_rad = float(accept_rad)
    o.analyser_darwin_hw_rad = float(darwin_hw_rad)
    o.gm_pixel_eval = int(bool(gm_pixel_eval))
    return o


def _counts_arr(counts):
    c = np.ascontiguousarray(np.asarray(counts, dtype=np.int64))
    return c, c.ctypes.data_as(ctypes.c_void_p)


def workspace_bytes(n_atoms, counts, n_frames, Ny, Nz):
    c, cp_ = _counts_arr(counts)
    out = ctypes.c_size_t(0)
    check(_lib.lib().mosaic_kin_workspace(int(n_atoms), int(c.shape[0]), cp_, int(n_frames),
                                          int(Ny), int(Nz), ctypes.byref(out)),
          "mosaic_kin_workspace")
    return int(out.value)


class StagedSample:
    """A sample staged into tiles inside a caller-owned workspace.

    Its statistics (``stage_stats``, ``radius_wl``, and ``pixel_geom`` once
    the pixel prologue ran) may stay in the workspace: ``kin_sum`` reads
    them there with a single wait, and the first read from Python fetches
    them (``mosaic_kin_stats``, which waits for the stream).
    """

    def __init__(self, n_atoms, counts, n_frames, Ny, Nz, workspace, stats, k, stream=None):
        self.n_atoms = int(n_atoms)
        self.k = float(k)
        self.counts = np.asarray(counts, dtype=np.int64)
        self.n_frames = int(n_frames)
        self.Ny, self.Nz = int(Ny), int(Nz)
        self.workspace = workspace
        self._stats = None if stats is None else (ctypes.c_double * 2)(*stats)
        self._pixel_geom = None
        self.pixels_key = None
        self.path_used = None
        self._stream = stream

    def _fetch(self, stage, pixels):
        ss = (ctypes.c_double * 2)() if stage else None
        pg = (ctypes.c_double * 8)() if pixels else None
        c, cp_ = _counts_arr(self.counts)
        check(_lib.lib().mosaic_kin_stats(self.n_atoms, int(c.shape[0]), cp_, self.n_frames,
                                          self.Ny, self.Nz, _ptr(self.workspace),
                                          int(self.workspace.numel()), ss, pg,
                                          _stream_handle(self._stream)),
              "mosaic_kin_stats")
        if stage:
            self._stats = ss
        if pixels:
            self._pixel_geom = pg

    @property
    def stage_stats(self):
        if self._stats is None:
            self._fetch(True, False)
        return self._stats

    @stage_stats.setter
    def stage_stats(self, v):
        self._stats = v

    @property
    def radius_wl(self):
        return float(self.stage_stats[0])

    @property
    def pixel_geom(self):
        """The pixel prologue's geometry (None before it ran on this workspace)."""
        if self._pixel_geom is None and self.pixels_key is not None:
            self._fetch(False, True)
        return self._pixel_geom

    @pixel_geom.setter
    def pixel_geom(self, v):
        self._pixel_geom = v


def ensure_workspace(ws, nbytes, device):
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
    return ws


def stage(pos, codes, amp, counts, k, n_frames, Ny, Nz, workspace=None, stream=None,
          readback=False):
    """Stage (N,3) float32 Angstrom sample-frame positions on the device.

    ``codes`` are uint8 species codes on the device, ``counts[s]`` the number
    of atoms with code s (host), ``amp`` an optional (N,) complex64 device
    tensor of incident amplitudes.  Returns without waiting for the device
    unless ``readback``: the statistics stay in the workspace until read.
    """
    if not pos.is_cuda:
        raise ValueError("positions must be a device tensor")
    if pos.dtype != torch.float32 or pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError("positions must be (N, 3) float32")
    pos = pos.contiguous()
    codes = codes.contiguous()
    n = int(pos.shape[0])
    if codes.dtype != torch.uint8 or codes.shape[0] != n:
        raise ValueError("species codes must be (N,) uint8")
    if amp is not None:
        if amp.dtype != torch.complex64 or amp.shape[0] != n:
            raise ValueError("amplitudes must be (N,) complex64")
        amp = torch.view_as_real(amp.contiguous())
    c, cp_ = _counts_arr(counts)
    if int(c.sum()) != n:
        raise ValueError("species counts do not add up to the atom count")
    nbytes = workspace_bytes(n, c, n_frames, Ny, Nz)
    ws = ensure_workspace(workspace, nbytes, pos.device)
    stats = (ctypes.c_double * 2)() if readback else None
    check(_lib.lib().mosaic_kin_stage(n, _ptr(pos), _ptr(codes), _ptr(amp), int(c.shape[0]),
                                      cp_, float(k), int(n_frames), int(Ny), int(Nz), _ptr(ws),
                                      int(ws.numel()), stats, _stream_handle(stream)),
          "mosaic_kin_stage")
    return StagedSample(n, c, n_frames, Ny, Nz, ws, None if stats is None else list(stats), k,
                        stream)


def restage(staged, pos, amp=None, stream=None):
    """Rebuild the tiles of ``staged`` from moved positions (same atoms, same
    species codes and order as its last ``stage``) without sorting; updates
    its stage statistics in place and returns it."""
    if not pos.is_cuda or pos.dtype != torch.float32 or tuple(pos.shape) != (staged.n_atoms, 3):
        raise ValueError("positions must be a (N, 3) float32 device tensor of the staged atoms")
    pos = pos.contiguous()
    if amp is not None:
        if amp.dtype != torch.complex64 or amp.shape[0] != staged.n_atoms:
            raise ValueError("amplitudes must be (N,) complex64")
        amp = torch.view_as_real(amp.contiguous())
    c, cp_ = _counts_arr(staged.counts)
    stats = (ctypes.c_double * 2)()
    check(_lib.lib().mosaic_kin_restage(staged.n_atoms, _ptr(pos), _ptr(amp), int(c.shape[0]),
                                        cp_, float(staged.k), staged.n_frames, staged.Ny,
                                        staged.Nz, _ptr(staged.workspace),
                                        int(staged.workspace.numel()), stats,
                                        _stream_handle(stream)),
          "mosaic_kin_restage")
    staged.stage_stats = stats
    return staged


def kin_pixels(staged, species_tab, opts, coords, stream=None, readback=False):
    """Pixel prologue into the staged sample's workspace.  With ``readback``
    waits for it and returns its geometry (rho_min [m], max row step of the
    pixel direction, max row second difference, max deviation from the
    matrix-core patch model, ...); else returns None at once, the geometry
    left in the workspace (``staged.pixel_geom`` fetches it)."""
    npix = staged.Ny * staged.Nz
    if coords.dtype != torch.float64 or tuple(coords.shape) != (3, npix) or not coords.is_cuda:
        raise ValueError("coords must be a (3, Ny*Nz) float64 device tensor")
    coords = coords.contiguous()
    geom = (ctypes.c_double * 8)() if readback else None
    check(_lib.lib().mosaic_kin_pixels(len(species_tab), ctypes.cast(species_tab, ctypes.c_void_p),
                                       ctypes.byref(opts), staged.Ny, staged.Nz, _ptr(coords),
                                       _ptr(staged.workspace), int(staged.workspace.numel()),
                                       geom, _stream_handle(stream)),
          "mosaic_kin_pixels")
    staged.pixel_geom = geom
    staged.pixels_key = _pixels_key(species_tab, opts, coords)
    return None if geom is None else tuple(geom)


def _pixels_key(species_tab, opts, coords):
    """What the pixel prologue's tables depend on: the coordinate tensor
    itself (held weakly, with its address and version counter, since a new
    tensor can be allocated at a freed one's address), species and flags."""
    return (weakref.ref(coords), int(coords.data_ptr()), int(coords._version),
            bytes(species_tab), bytes(opts))


def _same_pixels(key, species_tab, opts, coords):
    if key is None:
        return False
    ref, ptr, ver, sp, op = key
    return (ref() is coords and ptr == int(coords.data_ptr()) and ver == int(coords._version)
            and sp == bytes(species_tab) and op == bytes(opts))


def kin_sum(staged, species_tab, opts, rot, trans, coords, out, stream=None):
    """Accumulate the kinematic field of a staged sample into ``out``.

    rot (F,3,3) and trans (F,3) are float32 host arrays (sample -> lab,
    Angstrom), F at most the poses ``staged`` was staged for; coords a
    (3, Ny*Nz) float64 device tensor in metres; out a (F, Nz, Ny) complex64
    or complex128 device tensor.  Runs the pixel
    prologue first unless it already ran for these pixels, species and flags
    on this workspace.
    """
    rot = np.ascontiguousarray(np.asarray(rot, dtype=np.float32).reshape(-1, 3, 3))
    F = rot.shape[0]
    if F > staged.n_frames:
        raise ValueError(f"{F} poses for a workspace staged for {staged.n_frames}")
    trans = np.ascontiguousarray(np.asarray(trans, dtype=np.float32).reshape(F, 3))
    npix = staged.Ny * staged.Nz
    if coords.dtype != torch.float64 or tuple(coords.shape) != (3, npix) or not coords.is_cuda:
        raise ValueError("coords must be a (3, Ny*Nz) float64 device tensor")
    coords = coords.contiguous()
    if out.dtype not in (torch.complex64, torch.complex128) or out.numel() != F * npix:
        raise ValueError("out must hold (frames, Nz, Ny) complex64/complex128")
    if not out.is_contiguous():
        raise ValueError("out must be contiguous")
    if not _same_pixels(staged.pixels_key, species_tab, opts, coords):
        kin_pixels(staged, species_tab, opts, coords, stream)
    c, cp_ = _counts_arr(staged.counts)
    used = ctypes.c_int(0)
    check(_lib.lib().mosaic_kin_sum(
        staged.n_atoms, int(c.shape[0]), cp_, ctypes.cast(species_tab, ctypes.c_void_p),
        ctypes.byref(opts), F, rot.ctypes.data_as(ctypes.c_void_p),
        trans.ctypes.data_as(ctypes.c_void_p), staged.Ny, staged.Nz, _ptr(coords), _ptr(out),
        int(out.dtype == torch.complex128), staged._stats, staged._pixel_geom,
        _ptr(staged.workspace), int(staged.workspace.numel()), ctypes.byref(used),
        _stream_handle(stream)), "mosaic_kin_sum")
    staged.path_used = used.value
    return out


def sum_timing(enable=True, fallbacks=False):
    """Per-kernel launch times of the matrix-core moment sum (path "gemm") on
    the current device.  Switching timing on starts a fresh record of HIP
    events around its moments and evaluation kernels (up to 16 calls, nothing
    waits for them); each call returns the recorded calls' mean (moments,
    evaluation) times in milliseconds, waiting for their events, and with
    ``fallbacks`` also the number of (unit, patch) evaluations its separable
    phase model declined over those calls (evaluated pixel by pixel); then
    timing stays on or goes off by ``enable``."""
    ms = (ctypes.c_float * 3)()
    check(_lib.lib().mosaic_kin_timing(int(bool(enable)), ms), "mosaic_kin_timing")
    if fallbacks:
        return float(ms[0]), float(ms[1]), int(ms[2])
    return float(ms[0]), float(ms[1])


def ein_bilinear(pos, rot, trans, E0, du, dv, uc, vc, out=None, stream=None):
    """Incident amplitude per atom sampled from the beam grid E0 (NyB, NzB)."""
    n = int(pos.shape[0])
    if out is None:
        out = torch.empty(n, dtype=torch.complex64, device=pos.device)
    rot = np.ascontiguousarray(np.asarray(rot, dtype=np.float32).reshape(3, 3))
    trans = np.ascontiguousarray(np.asarray(trans, dtype=np.float32).reshape(3))
    E0 = E0.contiguous()
    NyB, NzB = int(E0.shape[0]), int(E0.shape[1])
    check(_lib.lib().mosaic_ein_bilinear(
        n, _ptr(pos.contiguous()), rot.ctypes.data_as(ctypes.c_void_p),
        trans.ctypes.data_as(ctypes.c_void_p), _ptr(torch.view_as_real(E0)), NyB, NzB,
        float(du), float(dv), float(uc), float(vc), _ptr(torch.view_as_real(out)),
        _stream_handle(stream)), "mosaic_ein_bilinear")
    return out


class DepthEin:
    """The depth-dependent entrance field of one sample on one beam grid
    (mosaic_ein_depth_*): the depth window, the beam-column map and the
    per-atom sampler, with their workspace."""

    def __init__(self, NyB, NzB, device):
        self.NyB, self.NzB = int(NyB), int(NzB)
        n = ctypes.c_size_t(0)
        check(_lib.lib().mosaic_ein_depth_workspace(self.NyB, self.NzB, ctypes.byref(n)),
              "mosaic_ein_depth_workspace")
        self.ws = torch.empty(max(1, n.value), dtype=torch.uint8, device=device)

    @staticmethod
    def _pose(rot, trans):
        r = np.ascontiguousarray(np.asarray(rot, dtype=np.float32).reshape(3, 3))
        t = np.ascontiguousarray(np.asarray(trans, dtype=np.float32).reshape(3))
        return r, t

    def bounds(self, pos, rot, trans, stream=None):
        """(s_min, s_max) along the beam [Angstrom]; synchronizes."""
        r, t = self._pose(rot, trans)
        out = (ctypes.c_double * 2)()
        check(_lib.lib().mosaic_ein_depth_bounds(
            int(pos.shape[0]), _ptr(pos.contiguous()), r.ctypes.data_as(ctypes.c_void_p),
            t.ctypes.data_as(ctypes.c_void_p), self.NyB, self.NzB, out, _ptr(self.ws),
            int(self.ws.numel()), _stream_handle(stream)), "mosaic_ein_depth_bounds")
        return float(out[0]), float(out[1])

    def column_map(self, pos, codes, fwd, rot, trans, du, dv, uc, vc, wavelength_A,
                   kernel_radius=0, A_out=None, stream=None):
        """Deposit the atoms' forward factors (fwd: (S, 2) float32 host,
        f0(0) + f' and f'') and form the sampled maps; optionally A (NyB, NzB)
        complex64 into ``A_out``."""
        r, t = self._pose(rot, trans)
        f = np.ascontiguousarray(np.asarray(fwd, dtype=np.float32).reshape(-1, 2))
        check(_lib.lib().mosaic_ein_column_map(
            int(pos.shape[0]), _ptr(pos.contiguous()), _ptr(codes.contiguous()), int(f.shape[0]),
            f.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p),
            t.ctypes.data_as(ctypes.c_void_p), self.NyB, self.NzB, float(du), float(dv),
            float(uc), float(vc), float(wavelength_A), int(kernel_radius),
            _ptr(torch.view_as_real(A_out) if A_out is not None else None), _ptr(self.ws),
            int(self.ws.numel()), _stream_handle(stream)), "mosaic_ein_column_map")

    def sample(self, pos, rot, trans, E0, du, dv, uc, vc, s_min, s_max, out=None, stream=None):
        """Per-atom Ein (N,) complex64 from the maps of the last column_map."""
        n = int(pos.shape[0])
        if out is None:
            out = torch.empty(n, dtype=torch.complex64, device=pos.device)
        r, t = self._pose(rot, trans)
        check(_lib.lib().mosaic_ein_depth(
            n, _ptr(pos.contiguous()), r.ctypes.data_as(ctypes.c_void_p),
            t.ctypes.data_as(ctypes.c_void_p), _ptr(torch.view_as_real(E0.contiguous())),
            self.NyB, self.NzB, float(du), float(dv), float(uc), float(vc), float(s_min),
            float(s_max), _ptr(torch.view_as_real(out)), _ptr(self.ws), int(self.ws.numel()),
            _stream_handle(stream)), "mosaic_ein_depth")
        return out


def mbeam_amplitude(pos, rot, trans, E, g, out=None, stream=None):
    """M-channel incident amplitude sum_m E[:, m] e^{2 pi i g_m . r} (N,)
    complex64; E (N, M) complex64 device, g (M, 3) host cycles per metre,
    r the lab position in metres (Beam.py:2283-2309)."""
    n = int(pos.shape[0])
    E = E.to(torch.complex64).contiguous()
    M = int(E.shape[1])
    gg = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(M, 3))
    if out is None:
        out = torch.empty(n, dtype=torch.complex64, device=pos.device)
    r = np.ascontiguousarray(np.asarray(rot, dtype=np.float32).reshape(3, 3))
    t = np.ascontiguousarray(np.asarray(trans, dtype=np.float32).reshape(3))
    check(_lib.lib().mosaic_mbeam_amplitude(
        n, _ptr(pos.contiguous()), r.ctypes.data_as(ctypes.c_void_p),
        t.ctypes.data_as(ctypes.c_void_p), _ptr(torch.view_as_real(E)), M,
        gg.ctypes.data_as(ctypes.c_void_p), _ptr(torch.view_as_real(out)),
        _stream_handle(stream)), "mosaic_mbeam_amplitude")
    return out


def same_amplitudes(a, b, rtol=1e-6):
    """Whether two poses' incident amplitudes agree to ``rtol`` of the largest.

    Inside a uniform beam the bilinear sampler returns 1 up to one float32
    rounding of its weights, which differs between poses; a shared amplitude
    then changes no atom's term by more than ~1e-7."""
    if a.numel() == 0:
        return True
    return bool(torch.max(torch.abs(a - b)) <= rtol * torch.max(torch.abs(a)))


def kin_adjoint(staged, k, rot, trans, grad_field, grad_pos, path="auto", stream=None,
                share_units=True):
    """Add dL/du (per Angstrom, original atom order) into grad_pos (N, 3) float32.

    ``grad_field`` is dL/dRe(A) + i dL/dIm(A) for the field the last
    ``kin_sum`` on ``staged`` produced, (F, Nz, Ny) complex (cast to
    complex64).  ``path`` is "auto", "rows", "mma" or "poly" (moment
    expansion); raises MosaicDeclined outside the chosen form's bounds.
    ``share_units`` lets the moment form compute each unit's pixel moments
    once (a scratch kept with ``staged``); False runs it tile by tile.
    """
    F = staged.n_frames
    if staged.pixel_geom is None:
        raise RuntimeError("kin_adjoint needs the pixel tables of a preceding kin_sum")
    rot = np.ascontiguousarray(np.asarray(rot, dtype=np.float32).reshape(F, 3, 3))
    trans = np.ascontiguousarray(np.asarray(trans, dtype=np.float32).reshape(F, 3))
    g = grad_field.to(torch.complex64).contiguous()
    if g.numel() != F * staged.Ny * staged.Nz:
        raise ValueError("grad_field must hold (frames, Nz, Ny) values")
    if grad_pos.dtype != torch.float32 or tuple(grad_pos.shape) != (staged.n_atoms, 3):
        raise ValueError("grad_pos must be (N, 3) float32")
    if not grad_pos.is_contiguous():
        raise ValueError("grad_pos must be contiguous")
    c, cp_ = _counts_arr(staged.counts)
    # scratch for the moment form's per-unit pixel moments, kept with the sample
    nbytes = ctypes.c_size_t(0)
    check(_lib.lib().mosaic_kin_adjoint_scratch(staged.n_atoms, int(c.shape[0]), cp_, F,
                                                staged.Ny, staged.Nz, ctypes.byref(nbytes)),
          "mosaic_kin_adjoint_scratch")
    scratch = getattr(staged, "adj_scratch", None)
    if not share_units:
        scratch = torch.empty(0, dtype=torch.uint8, device=staged.workspace.device)
    elif scratch is None or scratch.numel() < nbytes.value:
        scratch = torch.empty(int(nbytes.value), dtype=torch.uint8, device=staged.workspace.device)
        staged.adj_scratch = scratch
    check(_lib.lib().mosaic_kin_adjoint(
        staged.n_atoms, int(c.shape[0]), cp_, float(k), F, rot.ctypes.data_as(ctypes.c_void_p),
        trans.ctypes.data_as(ctypes.c_void_p), staged.Ny, staged.Nz, staged.stage_stats,
        staged.pixel_geom, _ptr(torch.view_as_real(g)), _ptr(grad_pos),
        _ptr(staged.workspace), int(staged.workspace.numel()), _ptr(scratch),
        int(scratch.numel()), PATHS[path], _stream_handle(stream)), "mosaic_kin_adjoint")
    return grad_pos


def scatter_workspace_bytes(n_atoms, Ny, Nz):
    out = ctypes.c_size_t(0)
    check(_lib.lib().mosaic_scatter_workspace(int(n_atoms), int(Ny), int(Nz), ctypes.byref(out)),
          "mosaic_scatter_workspace")
    return int(out.value)


def scatter(pos_m, f0p, f0z, sanr, sani, amp_r, amp_i, cx, cy, cz, k_val, remove_forward,
            Ny, Nz, apply_pol, pol_deg, pol_e, apply_decay, out_r, out_i, workspace=None,
            stream=None):
    """compute_scattering_cffi on the device (mosaic_scatter): the
    reference's arguments as device tensors -- positions (N, 3) float64 lab
    metres, WK rows (N, 11) float32, f0(0), f', f'', amplitude re/im (N,)
    float32, pixel coordinates (Ny*Nz,) float64 metres -- accumulating into
    the float32 device tensors out_r / out_i (Ny*Nz,).  Returns the workspace
    (reusable)."""
    n = int(pos_m.shape[0])
    arrs = [(pos_m, torch.float64, (n, 3)), (f0p, torch.float32, (n, 11)),
            (f0z, torch.float32, (n,)), (sanr, torch.float32, (n,)), (sani, torch.float32, (n,)),
            (amp_r, torch.float32, (n,)), (amp_i, torch.float32, (n,)),
            (cx, torch.float64, (Ny * Nz,)), (cy, torch.float64, (Ny * Nz,)),
            (cz, torch.float64, (Ny * Nz,)), (out_r, torch.float32, (Ny * Nz,)),
            (out_i, torch.float32, (Ny * Nz,))]
    for t, dt, shp in arrs:
        if not t.is_cuda or t.dtype != dt or tuple(t.shape) != shp or not t.is_contiguous():
            raise ValueError(f"mosaic_scatter: expected a contiguous {dt} device tensor {shp}")
    nbytes = scatter_workspace_bytes(n, Ny, Nz)
    ws = ensure_workspace(workspace, nbytes, out_r.device)
    check(_lib.lib().mosaic_scatter(
        n, _ptr(pos_m), _ptr(f0p), _ptr(f0z), int(bool(remove_forward)), _ptr(sanr), _ptr(sani),
        _ptr(amp_r), _ptr(amp_i), int(Ny), int(Nz), _ptr(cx), _ptr(cy), _ptr(cz), float(k_val),
        int(bool(apply_pol)), float(pol_deg), float(pol_e[0]), float(pol_e[1]), float(pol_e[2]),
        int(bool(apply_decay)), _ptr(out_r), _ptr(out_i), _ptr(ws), int(ws.numel()),
        _stream_handle(stream)), "mosaic_scatter")
    return ws
