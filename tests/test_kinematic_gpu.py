"""GPU parity of the kinematic sum against the CPU oracle.

Bar: relative complex-amplitude error ||A - A_ref|| / ||A_ref|| <= 1e-5
against the long-double oracle (the reference's FP64 kernel itself sits
~2e-6 from it on these inputs, from its own phase rounding), and the same
bound against the FP64 restatement.  Intensity error is reported as
max|I - I_ref| / max I_ref.
"""
import numpy as np
import pytest
import torch

from mosaic_amd import kinematic as K
from mosaic_amd import workloads as W
from mosaic_amd._lib import MosaicDeclined
from oracle import kinematic as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gpu_field(w, remove_forward=False, apply_pol=True, decay=False, path="auto"):
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
    I, IR = np.abs(A) ** 2, np.abs(R) ** 2
    e = rel(A, R)
    ei = float(np.max(np.abs(I - IR)) / np.max(IR))
    print(f"[parity] {name}: rel={e:.3e} int={ei:.3e}", flush=True)
    return e, ei


@pytest.mark.parametrize("path", ["tile", "rows", "mma", "poly"])
@pytest.mark.parametrize("flags", [
    dict(apply_pol=True, decay=False, remove_forward=False),
    dict(apply_pol=False, decay=True, remove_forward=False),
    dict(apply_pol=True, decay=True, remove_forward=True),
])
def test_fcc_small_vs_oracle(device, flags, path):
    w = W.fcc_au_bragg111(n=8, npix=32, seed=3)
    A = gpu_field(w, path=path, **flags)
    R80 = oracle_field(w, "f80", **flags)
    R64 = oracle_field(w, "f64", **flags)
    e80, _ = report(f"fcc8 {path} {flags}", A, R80)
    e64, _ = report(f"fcc8 {path} {flags} vs f64", A, R64)
    assert e80 <= TOL and e64 <= TOL


@pytest.mark.parametrize("path", ["tile", "rows", "mma", "poly"])
def test_perovskite_multispecies_chunks(device, path):
    # 24 pixels per row: the 8-pixel row blocks tile it exactly; 21 leaves a
    # ragged block at the row end
    w = W.perovskite_screw_bragg200(n=8, npix=21, blocks=(2, 2, 2))
    A = gpu_field(w, path=path)
    R = oracle_field(w, "f80")
    e, _ = report(f"perovskite8 {path} 3 species 8 chunks", A, R)
    assert e <= TOL


def test_gaussian_beam_complex_amplitude(device):
    """Incident amplitude varies per atom (Gaussian profile, small beam)."""
    w = W.fcc_au_bragg111(n=8, npix=24, seed=1)
    b = w["beam"]
    b.create_beam(10000.0, beam_size=(40.0, 40.0), beam_samples=(64, 64),
                  beam_profile="gaussian", pol_perp_rate=1.0)
    E0 = b._beam_E0_map * np.exp(1j * np.linspace(0, 2, 64))[None, :].astype(np.complex64)
    b.set_wavefield(E0.astype(np.complex64))
    A = gpu_field(w)
    # incident amplitudes from the oracle's restatement of the bilinear sampler
    pos = np.concatenate([w["sample"].load_chunk_positions(c) for c in
                          range(1, w["sample"].chunk_total + 1)])
    amp = O.ein_bilinear_numpy(pos, b._beam_E0_map, b._beam_du, b._beam_dv, b._beam_uc,
                               b._beam_vc)
    assert np.any(amp == 0) and np.any(np.abs(amp) > 0.5)
    R = oracle_field(w, "f80", amp=amp)
    e, _ = report("fcc8 gaussian complex beam", A, R)
    assert e <= TOL


def test_ein_bilinear_bitexact(device):
    rng = np.random.default_rng(5)
    pos = (rng.random((5000, 3)) * 80 - 40).astype(np.float32)
    E0 = (rng.random((33, 17)) + 1j * rng.random((33, 17))).astype(np.complex64)
    du, dv, uc, vc = 60.0 / 33, 50.0 / 17, 16.0, 8.0
    d = torch.from_numpy(pos).to(device)
    got = K.ein_bilinear(d, np.eye(3, dtype=np.float32), np.zeros(3, np.float32),
                         torch.from_numpy(E0).to(device), du, dv, uc, vc).cpu().numpy()
    want = O.ein_bilinear_numpy(pos, E0, du, dv, uc, vc)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("path", ["tile", "rows", "mma", "poly"])
def test_scan_frames_match_single_poses(device, path):
    w = W.fcc_fault_scan111(cells=(6, 6, 8), npix=24, n_angles=5, half_range_deg=0.5)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    b._kin_path = path
    frames = b.kinematic_scan(s, d, st, "phi", w["angles_deg"], apply_polarization=True,
                              out_dtype=torch.complex128).cpu().numpy()
    for i in (0, 2, 4):
        st.set_single_motor_value_absolute("phi", float(w["angles_deg"][i]))
        R = oracle_field(w, "f80")
        e, _ = report(f"scan {path} frame {i}", frames[i].ravel(), R)
        assert e <= TOL
    st.set_single_motor_value_absolute("phi", 0.0)


@pytest.mark.parametrize("path", ["mma", "poly"])
@pytest.mark.parametrize("shape", [(1, 37), (37, 1), (3, 2), (17, 33), (70, 45)])
def test_patch_kernels_ragged(device, shape, path):
    """Detectors whose sides are not multiples of the 32-pixel patch, down to
    single rows and columns (one- and two-point patch models), for the
    matrix-core and the moment sums."""
    w = W.fcc_au_bragg111(n=6, npix=8, seed=4)
    b, d = w["beam"], w["detector"]
    d.create_detector(shape, (7.5e5, 7.5e5))
    d.position_detector_absolute(1.0e10, np.rad2deg(2 * w["theta_b"]), 0.0)
    A = gpu_field(w, path=path)
    assert b._last_path % 100 in ((32,) if path == "mma" else (45, 46))
    R = oracle_field(w, "f80")
    e, _ = report(f"{path} detector {shape}", A, R)
    assert e <= TOL


def test_kd_tiles_are_compact(device):
    """The k-d bisection keeps every 256-atom tile near the radius of a cube
    of 256 atoms (17 wavelengths for the 32,000-atom crystal, 21 for the
    perovskite's sparsest species; a Morton order gives 46 and 95)."""
    for w, bound in ((W.fcc_au_bragg111(n=20, npix=8), 20.0),
                     (W.perovskite_screw_bragg200(n=24, npix=8, blocks=(2, 2, 2)), 25.0)):
        b, d = w["beam"], w["detector"]
        pos, codes, counts = [], [], None
        smp = w["sample"]
        for cid in range(1, smp.chunk_total + 1):
            pos.append(smp.load_chunk_positions(cid))
            codes.append(smp.load_chunk_species(cid, codes=True)[0])
        pos = np.concatenate(pos).astype(np.float32)
        codes = np.concatenate(codes).astype(np.uint8)
        order = np.argsort(codes, kind="stable")
        pos, codes = pos[order], codes[order]
        counts = np.bincount(codes).astype(np.int64)
        k = 2 * np.pi / b.wavelength
        st = K.stage(torch.as_tensor(pos, device=device), torch.as_tensor(codes, device=device),
                     None, counts, k, 1, 8, 8)
        print(f"[kd] {pos.shape[0]} atoms: largest tile radius {st.radius_wl:.1f} wavelengths")
        assert 0.0 < st.radius_wl <= bound


def test_restage_moved_atoms(device):
    """Tiles rebuilt from displaced positions with the original k-d order
    give the field of the displaced crystal (long-double oracle), with tile
    radii close to a fresh staging's."""
    from mosaic_amd import tables
    w = W.fcc_au_bragg111(n=10, npix=24, seed=5)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    pos0 = s.load_chunk_positions(1).astype(np.float32)
    rng = np.random.default_rng(9)
    pos1 = (pos0 + W.gaussian_bump_field(pos0, 0.8, 8.0, rng)).astype(np.float32)
    k = 2 * np.pi / b.wavelength
    counts = np.array([pos0.shape[0]], np.int64)
    codes = torch.zeros(pos0.shape[0], dtype=torch.uint8, device=device)
    Ny, Nz = d.shape
    spc = K.species_table(tables.species_constants(["Au"], b.energy))
    opts = K.kin_opts(k, False, True, b._polarization_state(), False)
    coords = torch.as_tensor(np.ascontiguousarray(d.pixel_coordinates / 1e10), device=device)
    rot = np.asarray(st.rotation, np.float32)[None]
    tr = np.asarray(st.translation, np.float32)[None]
    staged = K.stage(torch.as_tensor(pos0, device=device), codes, None, counts, k, 1, Ny, Nz)
    r0 = staged.radius_wl
    K.restage(staged, torch.as_tensor(pos1, device=device))
    out = torch.zeros((1, Nz, Ny), dtype=torch.complex128, device=device)
    K.kin_sum(staged, spc, opts, rot, tr, coords, out)
    fresh = K.stage(torch.as_tensor(pos1, device=device), codes, None, counts, k, 1, Ny, Nz)
    assert staged.radius_wl <= fresh.radius_wl + 2.0 and staged.radius_wl >= r0 - 1.0
    from mosaic_amd.sample import sample
    w["sample"] = sample([pos1], [np.zeros(pos1.shape[0], np.uint8)], ["Au"])
    R = oracle_field(w, "f80")
    e, _ = report("restaged displaced crystal", out[0].cpu().numpy().ravel(), R)
    assert e <= TOL


def test_statistics_left_in_workspace(device):
    """Staging and the pixel prologue without read-back (the sum reads their
    statistics from the workspace, one wait) give the same field, bit for
    bit, as with the statistics read back to the host, and the statistics
    fetched afterwards equal those read back."""
    from mosaic_amd import tables
    w = W.perovskite_screw_bragg200(n=10, npix=64, blocks=(2, 2, 2))
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    pos = torch.as_tensor(np.ascontiguousarray(s.load_chunk_positions(1), np.float32),
                          device=device)
    codes_h = np.asarray(s.load_chunk_species(1, codes=True)[0], np.uint8)
    codes = torch.as_tensor(codes_h, device=device)
    counts = np.bincount(codes_h, minlength=3).astype(np.int64)
    k = 2 * np.pi / b.wavelength
    Ny, Nz = d.shape
    spc = K.species_table(tables.species_constants(["Sr", "Ti", "O"], b.energy))
    opts = K.kin_opts(k, False, True, b._polarization_state(), False)
    coords = torch.as_tensor(np.ascontiguousarray(d.pixel_coordinates / 1e10), device=device)
    rot = np.asarray(st.rotation, np.float32)[None]
    tr = np.asarray(st.translation, np.float32)[None]
    fields, stats = [], []
    for readback in (True, False):
        staged = K.stage(pos, codes, None, counts, k, 1, Ny, Nz, readback=readback)
        K.kin_pixels(staged, spc, opts, coords, readback=readback)
        out = torch.zeros((1, Nz, Ny), dtype=torch.complex64, device=device)
        K.kin_sum(staged, spc, opts, rot, tr, coords, out)
        fields.append(out.cpu().numpy())
        stats.append((tuple(staged.stage_stats), tuple(staged.pixel_geom), staged.path_used))
    assert stats[0] == stats[1]
    assert np.array_equal(fields[0], fields[1])


def test_restage_keeps_units(device):
    """Restaging the staged positions rebuilds the same tiles and re-measures
    the radii of the pairs and quads of tiles, so the moment sum forms the
    same units and the image is unchanged; summing each tile alone instead
    differs by the expansion error (about 1e-6)."""
    from mosaic_amd import tables
    w = W.fcc_au_bragg111(n=10, npix=24, seed=5)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    pos = torch.as_tensor(s.load_chunk_positions(1).astype(np.float32), device=device)
    k = 2 * np.pi / b.wavelength
    counts = np.array([pos.shape[0]], np.int64)
    codes = torch.zeros(pos.shape[0], dtype=torch.uint8, device=device)
    Ny, Nz = d.shape
    spc = K.species_table(tables.species_constants(["Au"], b.energy))
    opts = K.kin_opts(k, False, True, b._polarization_state(), False)
    coords = torch.as_tensor(np.ascontiguousarray(d.pixel_coordinates / 1e10), device=device)
    rot = np.asarray(st.rotation, np.float32)[None]
    tr = np.asarray(st.translation, np.float32)[None]
    staged = K.stage(pos, codes, None, counts, k, 1, Ny, Nz)
    a = torch.zeros((1, Nz, Ny), dtype=torch.complex128, device=device)
    K.kin_sum(staged, spc, opts, rot, tr, coords, a)
    path_a = staged.path_used
    K.restage(staged, pos)
    b2 = torch.zeros_like(a)
    K.kin_sum(staged, spc, opts, rot, tr, coords, b2)
    path_b = staged.path_used
    diff = float((a - b2).abs().max() / a.abs().max())
    print(f"[restage] same positions: path {path_a} -> {path_b}, max diff {diff:.3e}")
    assert path_a == path_b and path_a // 10 % 10 == 4
    assert diff <= 1e-9


def test_near_field_forced_forms_decline_auto_is_exact(device):
    """At 1 mm the far-field and per-tile forms decline when forced (nothing
    written); the automatic choice runs the exact per-pair kernel, which
    matches the long-double oracle."""
    w = W.fcc_au_bragg111(n=4, npix=8)
    d = w["detector"]
    d.create_detector((8, 8), (750.0, 750.0))
    d.position_detector_absolute(1e7, np.rad2deg(2 * w["theta_b"]), 0.0)  # 1 mm
    for path in ("tile", "rows", "mma", "poly"):
        with pytest.raises(MosaicDeclined):
            gpu_field(w, path=path)
    A = gpu_field(w, path="auto")
    assert w["beam"]._last_path == 50
    e, _ = report("near field 1 mm auto", A, oracle_field(w, "f80"))
    assert e <= TOL


def test_empty_and_ragged(device):
    from mosaic_amd.sample import sample
    w = W.fcc_au_bragg111(n=6, npix=16, sigma=0.0)
    pos = w["sample"].load_chunk_positions(1)
    # ragged chunks: 0, 1, 255, 257 atoms and the rest; two species
    cuts = [0, 0, 1, 256, 513, pos.shape[0]]
    chunks = [pos[cuts[i]:cuts[i + 1]] for i in range(len(cuts) - 1)]
    codes = [np.arange(len(c), dtype=np.uint8) % 2 for c in chunks]
    w["sample"] = sample(chunks, codes, ["Au", "Cu"])
    A = gpu_field(w)
    R = oracle_field(w, "f80")
    e, _ = report("ragged chunks 2 species", A, R)
    assert e <= TOL
    w["sample"] = sample([pos[:0]], [np.zeros(0, np.uint8)], ["Au"])
    A0 = gpu_field(w)
    assert np.all(A0 == 0)


def test_c1_full_vs_fp64_oracle(device):
    """32,000 atoms x 128^2 pixels (5.2e8 pairs) against the FP64 restatement,
    all four kernels; the automatic choice is the moment sum here."""
    w = W.fcc_au_bragg111(n=20, npix=128, seed=0)
    R = oracle_field(w, "f64")
    for path in ("tile", "rows", "mma", "poly", "auto"):
        A = gpu_field(w, path=path)
        e, ei = report(f"C1 full {path} vs f64", A, R)
        assert e <= TOL and ei <= TOL
    assert w["beam"]._last_path % 100 in (45, 46)


def test_rows_second_order_close_detector(device):
    """A 5 cm detector with 10 um pixels: the tiles' own second-order phase
    (~3e-7 rev) is above the first-order bound, the rows are still straight
    enough for 4-pixel lanes and the patches for the quadratic model, so the
    recurrence and the matrix-core form run with the second-order term; all
    kernels match the long-double sum."""
    w = W.fcc_au_bragg111(n=8, npix=32, seed=2)
    d = w["detector"]
    d.create_detector((32, 32), (1.0e5, 1.0e5))
    d.position_detector_absolute(5.0e8, np.rad2deg(2 * w["theta_b"]), 0.0)
    R = oracle_field(w, "f80")
    want = {"rows": (124, 128), "mma": (132,), "tile": (12,)}
    for path in ("rows", "mma", "tile"):
        A = gpu_field(w, path=path)
        e, _ = report(f"5 cm detector {path} (path {w['beam']._last_path})", A, R)
        assert e <= TOL
        assert w["beam"]._last_path in want[path]


def test_auto_path_choice(device):
    """Far field picks the moment sum; a 5 cm detector with 2 mm pixels
    bends the rows and patches too much for it, the matrix-core form and the
    recurrence: its 256 pixels take the probe form (one lane per pixel), and
    with the probe excluded (forced per-tile path) the per-tile factors."""
    w = W.fcc_au_bragg111(n=4, npix=16)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    b._kin_path = "auto"
    gpu_field(w)
    staged_path = b._last_path
    assert staged_path in (45, 46)
    d.create_detector((16, 16), (2.0e7, 2.0e7))
    d.position_detector_absolute(5.0e8, np.rad2deg(2 * w["theta_b"]), 0.0)
    A = gpu_field(w)
    assert b._last_path in (60, 160)
    R = oracle_field(w, "f80")
    e, _ = report("near detector auto", A, R)
    assert e <= TOL
    A = gpu_field(w, path="tile")
    assert b._last_path in (11, 12)
    e, _ = report("near detector per-tile", A, R)
    assert e <= TOL


class _PixelBlock:
    """A rectangular block of another detector's pixels, at the same positions."""

    def __init__(self, det, y0, z0, ny, nz):
        Ny, Nz = det.shape
        c = np.asarray(det.pixel_coordinates).reshape(3, Nz, Ny)
        self.pixel_coordinates = np.ascontiguousarray(c[:, z0:z0 + nz, y0:y0 + ny].reshape(3, -1))
        self.shape = (ny, nz)


def test_moment_sum_cross_term_regime(device):
    """Corner pixels of a 512^2, 75 um detector at 1 m and an FCC crystal
    stretched 1.5 times with 0.5 A random displacements (largest tile ~28
    wavelengths; the disorder keeps the field diffuse, where a perfect
    stretched lattice would cancel to ~1e-3 of sqrt(N) f and leave the
    relative error ill-conditioned): the separable patch model's deviation
    times the radius exceeds 1e-6 (the matrix-core form declines), the model
    with the cross term a b g_yz stays below it, and the moment sum matches
    the long-double oracle."""
    from mosaic_amd import sample as S
    w = W.fcc_au_bragg111(n=12, npix=8, sigma=0.0)
    rng = np.random.default_rng(3)
    base = w["sample"].load_chunk_positions(1)
    pos = (base * 1.5 + rng.normal(0.0, 0.5, base.shape)).astype(np.float32)
    w["sample"] = S.sample([pos], [np.zeros(pos.shape[0], np.uint8)], ["Au"])
    w["detector"] = _PixelBlock(W._detector(512, 2 * w["theta_b"]), 0, 0, 64, 64)
    A = gpu_field(w)
    r, geom = w["beam"]._last_stats
    print(f"[cross] radius {r:.2f} wl, separable {geom[3] * r:.2e}, with cross term "
          f"{geom[6] * r:.2e}, path {w['beam']._last_path}", flush=True)
    assert geom[3] * r > 1e-6 >= geom[6] * r
    assert w["beam"]._last_path % 100 in (45, 46)
    # the stretched crystal samples the beam's amplitude map where it is not flat
    bm = w["beam"]
    amp = O.ein_bilinear_numpy(pos, bm._beam_E0_map, bm._beam_du, bm._beam_dv, bm._beam_uc,
                               bm._beam_vc)
    R = oracle_field(w, "f80", amp=amp)
    e, _ = report("cross-term regime (moment sum)", A, R)
    assert e <= TOL


def test_staging_many_species_ragged(device):
    """Random atoms of 12 species, interleaved in the input, with counts on
    and around the staging's boundaries: none, one atom, one tile +-1, the
    2,048-position chunk +-1, several global k-d levels for the largest
    species; the field against the long-double oracle."""
    from mosaic_amd.sample import sample
    w = W.fcc_au_bragg111(n=4, npix=16, sigma=0.0)
    labels = ["Au", "Cu", "Ag", "Pt", "Pd", "Ni", "Fe", "Co", "Al", "Si", "Ti", "Sr"]
    counts = [40000, 0, 1, 255, 256, 257, 2047, 2048, 2049, 4097, 513, 6000]
    n = sum(counts)
    rng = np.random.default_rng(11)
    side = (n * 17.0) ** (1.0 / 3.0)
    pos = ((rng.random((n, 3)) - 0.5) * side).astype(np.float32)
    codes = np.repeat(np.arange(len(counts), dtype=np.uint8), counts)
    perm = rng.permutation(n)
    pos, codes = np.ascontiguousarray(pos[perm]), np.ascontiguousarray(codes[perm])
    w["sample"] = sample([pos], [codes], labels)
    A = gpu_field(w)
    R = oracle_field(w, "f80")
    e, _ = report(f"{n} random atoms, 12 species, ragged counts (path {w['beam']._last_path})",
                  A, R)
    assert e <= TOL


def test_probe_scattered_pixels(device):
    """The probe form (one lane per frame and pixel) on 40 scattered pixels
    of a 32 x 32 detector at three poses, against the long-double oracle at
    those pixels, and against the moment sum over the whole detector."""
    w = W.fcc_au_bragg111(n=8, npix=32, seed=3)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    pc = np.asarray(d.pixel_coordinates)
    sel = np.sort(np.random.default_rng(4).choice(32 * 32, 40, replace=False))
    poses = []
    for a in (-0.02, 0.0, 0.03):
        st.set_single_motor_value_absolute("phi", a)
        poses.append((np.asarray(st.rotation, np.float32), np.asarray(st.translation, np.float32)))
    st.set_single_motor_value_absolute("phi", 0.0)
    rot = np.stack([q[0] for q in poses])
    tr = np.stack([q[1] for q in poses])
    b._kin_path = "probe"
    got = b._kin_field(s, np.ascontiguousarray(pc[:, sel]), (40, 1), st, (rot, tr), False,
                       True, False)
    torch.cuda.synchronize()
    assert b._last_path in (60, 160)
    b._kin_path = "auto"
    full = b._kin_field(s, pc, d.shape, st, (rot, tr), False, True, False)
    torch.cuda.synchronize()
    got = got.cpu().numpy().reshape(3, 40)
    full = full.cpu().numpy().reshape(3, -1)[:, sel]
    for i, (R, T) in enumerate(poses):
        a = W.oracle_inputs(w, frame_rot=R, frame_trans=T)
        kw = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i",
                                "k_val", "pol_deg", "pol_e")}
        ref = O.kin_sum_c(**kw, cx=a["cx"][sel], cy=a["cy"][sel], cz=a["cz"][sel], Ny=40, Nz=1,
                          remove_forward=False, apply_pol=True, apply_decay=False,
                          precision="f80")
        e, _ = report(f"probe pose {i}", got[i], ref)
        assert e <= TOL
        assert rel(got[i], full[i]) <= TOL
