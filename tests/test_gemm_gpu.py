"""GPU parity of the matrix-core moment sum (k_gm_moments + k_gm_eval, path
"gemm") against the CPU oracle and against the VALU moment sum.

Bar: relative complex-amplitude error <= 1e-5 against the long-double and
the FP64 oracles (as every sum kernel); the two moment sums, which share the
expansion's bounds, within 1e-5 of each other.
"""
import numpy as np
import pytest
import torch

from mosaic_amd import workloads as W
from oracle import kinematic as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
GEMM = (76, 176)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gpu_field(w, path="gemm", remove_forward=False, apply_pol=True, decay=False):
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    b._kin_path = path
    rot = np.asarray(st.rotation, dtype=np.float32)[None]
    tr = np.asarray(st.translation, dtype=np.float32)[None]
    out = b._kin_field(s, d.pixel_coordinates, d.shape, st, (rot, tr), remove_forward,
                       apply_pol, decay)
    torch.cuda.synchronize()
    return out[0].cpu().numpy().ravel()


def oracle_field(w, precision="f80", remove_forward=False, apply_pol=True, decay=False,
                 amp=None):
    a = W.oracle_inputs(w, amp=amp)
    kw = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx",
                            "cy", "cz", "k_val", "Ny", "Nz", "pol_deg", "pol_e")}
    return O.kin_sum_c(**kw, remove_forward=remove_forward, apply_pol=apply_pol,
                       apply_decay=decay, precision=precision)


def report(name, A, R):
    e = rel(A, R)
    ei = float(np.max(np.abs(np.abs(A) ** 2 - np.abs(R) ** 2)) / np.max(np.abs(R) ** 2))
    print(f"[gemm] {name}: rel={e:.3e} int={ei:.3e}", flush=True)
    return e


@pytest.mark.parametrize("flags", [
    dict(apply_pol=True, decay=False, remove_forward=False),
    dict(apply_pol=False, decay=True, remove_forward=False),
    dict(apply_pol=True, decay=True, remove_forward=True),
])
def test_gemm_small_vs_oracle(device, flags):
    w = W.fcc_au_bragg111(n=8, npix=32, seed=3)
    A = gpu_field(w, **flags)
    assert w["beam"]._last_path in GEMM
    e80 = report(f"fcc8 {flags}", A, oracle_field(w, "f80", **flags))
    e64 = report(f"fcc8 {flags} vs f64", A, oracle_field(w, "f64", **flags))
    assert e80 <= TOL and e64 <= TOL


def test_gemm_multispecies_chunks(device):
    w = W.perovskite_screw_bragg200(n=8, npix=21, blocks=(2, 2, 2))
    A = gpu_field(w)
    e = report("perovskite8 3 species 8 chunks", A, oracle_field(w, "f80"))
    assert e <= TOL


@pytest.mark.parametrize("shape", [(1, 37), (37, 1), (3, 2), (17, 33), (70, 45), (160, 130)])
def test_gemm_ragged_detectors(device, shape):
    """Detector sides that are not multiples of the 32-pixel patch, single
    rows and columns, and a region cut by the detector edge
    (160 x 130: 5 x 5 patches in one region, each of its four groups
    partial)."""
    w = W.fcc_au_bragg111(n=6, npix=8, seed=4)
    d = w["detector"]
    d.create_detector(shape, (7.5e5, 7.5e5))
    d.position_detector_absolute(1.0e10, np.rad2deg(2 * w["theta_b"]), 0.0)
    A = gpu_field(w)
    assert w["beam"]._last_path in GEMM
    e = report(f"detector {shape}", A, oracle_field(w, "f80"))
    assert e <= TOL


def test_gemm_c1_full(device):
    """32,000 atoms x 128^2 pixels (16 patches: one wave of a region live)
    against the FP64 restatement."""
    w = W.fcc_au_bragg111(n=20, npix=128, seed=0)
    A = gpu_field(w)
    assert w["beam"]._last_path in GEMM
    e = report("C1 full vs f64", A, oracle_field(w, "f64"))
    assert e <= TOL


def test_gemm_scan_frames(device):
    """All poses of a scan in one launch pair, each frame against the oracle."""
    w = W.fcc_fault_scan111(cells=(6, 6, 8), npix=24, n_angles=5, half_range_deg=0.5)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    b._kin_path = "gemm"
    frames = b.kinematic_scan(s, d, st, "phi", w["angles_deg"], apply_polarization=True,
                              out_dtype=torch.complex128).cpu().numpy()
    for i in (0, 2, 4):
        st.set_single_motor_value_absolute("phi", float(w["angles_deg"][i]))
        e = report(f"scan frame {i}", frames[i].ravel(), oracle_field(w, "f80"))
        assert e <= TOL
    st.set_single_motor_value_absolute("phi", 0.0)


class _PixelBlock:
    """A rectangular block of another detector's pixels, at the same positions."""

    def __init__(self, det, y0, z0, ny, nz):
        Ny, Nz = det.shape
        c = np.asarray(det.pixel_coordinates).reshape(3, Nz, Ny)
        self.pixel_coordinates = np.ascontiguousarray(c[:, z0:z0 + nz, y0:y0 + ny].reshape(3, -1))
        self.shape = (ny, nz)


def test_gemm_region_gradients_at_detector_corner(device):
    """The corner 256 x 256 pixels of a 512^2, 75 um detector at 1 m (one
    region of 8 x 8 patches, the largest change of the gradients across it)
    and an FCC crystal stretched 1.5 times with 0.5 A random displacements
    (largest tile ~28 wavelengths): the region's mean gradients differ from
    its patches' by a few 1e-3 rad over a patch, which the expansion carries
    to second order; the field against the long-double oracle."""
    from mosaic_amd import sample as S
    w = W.fcc_au_bragg111(n=12, npix=8, sigma=0.0)
    rng = np.random.default_rng(3)
    base = w["sample"].load_chunk_positions(1)
    pos = (base * 1.5 + rng.normal(0.0, 0.5, base.shape)).astype(np.float32)
    w["sample"] = S.sample([pos], [np.zeros(pos.shape[0], np.uint8)], ["Au"])
    w["detector"] = _PixelBlock(W._detector(512, 2 * w["theta_b"]), 0, 0, 256, 256)
    A = gpu_field(w)
    r, _ = w["beam"]._last_stats
    print(f"[gemm] corner block: radius {r:.2f} wl, path {w['beam']._last_path}", flush=True)
    assert w["beam"]._last_path in GEMM
    bm = w["beam"]
    amp = O.ein_bilinear_numpy(pos, bm._beam_E0_map, bm._beam_du, bm._beam_dv, bm._beam_uc,
                               bm._beam_vc)
    e = report("corner block, stretched crystal", A, oracle_field(w, "f80", amp=amp))
    assert e <= TOL


def test_gemm_second_order_close_detector(device):
    """A 10 cm detector of 256^2 pixels of 10 um: the tiles' own second-order
    phase is above the first-order bound, so the matrix-core moment sum runs
    with it (path 176, the phase products with the |s|^2 and (xc . s)^2
    terms); the field against the long-double oracle."""
    w = W.fcc_au_bragg111(n=8, npix=256, seed=2)
    d = w["detector"]
    d.create_detector((256, 256), (1.0e5, 1.0e5))
    d.position_detector_absolute(1.0e9, np.rad2deg(2 * w["theta_b"]), 0.0)
    A = gpu_field(w)
    assert w["beam"]._last_path == 176
    e = report("10 cm detector, second order", A, oracle_field(w, "f80"))
    assert e <= TOL


def test_gemm_units_across_slices(device):
    """A 36^3-cell perovskite (233,280 atoms; Sr and Ti each 183 tiles, the
    last of 64 atoms at an even place, so a unit of one tile and one chunk)
    on a 512^2 detector (four regions, slices of about 15 tiles): each
    unit's first chunk is built during the previous unit's last, and the
    field agrees with the VALU moment sum."""
    w = W.perovskite_screw_bragg200(n=36, npix=512)
    A = gpu_field(w, path="auto")
    assert w["beam"]._last_path in GEMM
    B = gpu_field(w, path="poly")
    assert w["beam"]._last_path % 100 in (45, 46)
    e = report("perovskite36 x 512^2: gemm vs poly", A, B)
    assert e <= TOL


def test_gemm_matches_valu_moment_sum(device):
    """A 200,000-atom FCC crystal on a 256^2 detector (one region of 8 x 8
    patches, units of up to four tiles): the automatic choice takes the
    matrix-core moment sum, and it agrees with the VALU moment sum."""
    w = W.fcc_au_bragg111(n=37, npix=256, seed=1)
    A = gpu_field(w, path="auto")
    assert w["beam"]._last_path in GEMM
    B = gpu_field(w, path="poly")
    assert w["beam"]._last_path % 100 in (45, 46)
    e = report("200k atoms x 256^2: gemm vs poly", A, B)
    assert e <= TOL


def _staged_field(w, pixel_eval, device):
    """The field of workload ``w`` through the staged C-ABI calls with path
    "gemm", and the number of (unit, patch) evaluations taken pixel by pixel."""
    from mosaic_amd import kinematic as K
    from mosaic_amd import tables
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    pos = s.load_chunk_positions(1).astype(np.float32)
    codes = np.zeros(pos.shape[0], np.uint8)
    k = 2.0 * np.pi / b.wavelength
    spc = K.species_table(tables.species_constants(["Au"], b.energy))
    opts = K.kin_opts(k, False, True, b._polarization_state(), False, path="gemm",
                      gm_pixel_eval=pixel_eval)
    rot = np.asarray(st.rotation, np.float32)[None]
    tr = np.asarray(st.translation, np.float32)[None]
    coords = torch.from_numpy(np.ascontiguousarray(d.pixel_coordinates / 1e10)).to(device)
    Ny, Nz = d.shape
    d_pos = torch.from_numpy(pos).to(device)
    amp = K.ein_bilinear(d_pos, rot[0], tr[0], torch.from_numpy(b._beam_E0_map).to(device),
                         b._beam_du, b._beam_dv, b._beam_uc, b._beam_vc)
    staged = K.stage(d_pos, torch.from_numpy(codes).to(device), amp, [pos.shape[0]], k, 1, Ny, Nz)
    out = torch.zeros((1, Nz, Ny), dtype=torch.complex128, device=device)
    K.sum_timing(True)
    try:
        K.kin_sum(staged, spc, opts, rot, tr, coords, out)
        torch.cuda.synchronize()
        _, _, n_fb = K.sum_timing(True, fallbacks=True)
    finally:
        K.sum_timing(False)
    return out[0].cpu().numpy().ravel(), staged.path_used, n_fb


def test_gemm_pixel_eval_fallback(device):
    """The evaluation's per-pixel fallback (taken for a unit whose origin is
    too far from its slice's reference for the separable phase model), forced
    for every unit: 32,000 atoms x 128^2 pixels (16 patches) against the FP64
    restatement and against the separable evaluation of the same moments;
    the count shows which evaluation ran (every (unit, patch), or none)."""
    w = W.fcc_au_bragg111(n=20, npix=128, seed=0)
    n_patches = 16
    A, path, n_fb = _staged_field(w, True, device)
    B, path_b, n_sep = _staged_field(w, False, device)
    print(f"[gemm] pixel evaluation: paths {path} {path_b}, per-pixel (unit, patch) "
          f"{n_fb} forced, {n_sep} by the model's check", flush=True)
    assert path in GEMM and path_b in GEMM
    n_tiles = (32000 + 255) // 256
    assert n_fb % n_patches == 0 and n_fb >= n_patches * ((n_tiles + 3) // 4)
    assert n_sep < n_fb
    e = report("pixel evaluation vs f64", A, oracle_field(w, "f64"))
    e2 = report("pixel evaluation vs separable", A, B)
    assert e <= TOL and e2 <= TOL
