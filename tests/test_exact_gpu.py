"""GPU parity of the exact per-pair kernels.

* ``mosaic_scatter`` (compute_scattering_cffi on the device), called through
  the reference-side ctypes stub (tests/cffi_stub.py), on every golden case of
  the reference's own C kernel: 4 geometries (including pixels 30 nm away and
  a forward corner where remove_forward cuts) x 5 flag sets x 2
  polarizations (sigma, and p = 0.3 along z).  Bar: the golden bar of the CPU
  oracle, max |A - A_gold| <= 5e-5 max |A_gold| (two float64 evaluations of
  k (x + r) ~ 5e10 rad differ by ~1e-5 rad per pair), and relative L2 <= 1e-5
  against the long-double oracle.
* the staged sum's exact path (``mosaic_kin_sum`` path 5), which the automatic
  choice takes wherever the other kernels' bounds fail: the same golden
  cases, near-field crystals, and a forced far-field run, against the oracle
  at 1e-5.
* the analyser (top-hat, Darwin) against closed forms for one atom at the
  origin and against the oracle's restatement (Beam.py:2350-2378).
"""
import numpy as np
import pytest
import torch

from mosaic_amd import _lib
from mosaic_amd import kinematic as K
from mosaic_amd import workloads as W
from oracle import kinematic as O

import cffi_stub

pytestmark = pytest.mark.gpu

GOLD = np.load(__file__.replace("test_exact_gpu.py", "golden/kin_golden.npz"))
CASES = ("three_atoms", "random64", "forward25", "near30nm")
TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def golden_args(name):
    g = GOLD
    pc = g[f"{name}/coords"]
    ny, nz = (int(v) for v in g[f"{name}/shape"])
    amp = g[f"{name}/amp"]
    return dict(pos_m=g[f"{name}/pos"], f0p=g[f"{name}/f0p"], f0z=g[f"{name}/f0z"],
                sanr=g[f"{name}/sanr"], sani=g[f"{name}/sani"],
                amp_r=amp.real.astype(np.float32), amp_i=amp.imag.astype(np.float32),
                cx=pc[0], cy=pc[1], cz=pc[2], k_val=float(g["k"]), Ny=ny, Nz=nz)


def flags(fi, pi):
    rf, pol, dec = (bool(v) for v in GOLD["flags"][fi])
    pv = GOLD["pol"][pi]
    return dict(remove_forward=rf, apply_pol=pol, apply_decay=dec, pol_deg=float(pv[0]),
                pol_e=tuple(float(v) for v in pv[1:]))


def stub_call(lib, a, f, device, repeats=1):
    dev = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(device)
    n = int(a["pos_m"].shape[0])
    npix = a["Ny"] * a["Nz"]
    t = [dev(a["pos_m"], np.float64), dev(a["f0p"], np.float32), dev(a["f0z"], np.float32),
         dev(a["sanr"], np.float32), dev(a["sani"], np.float32), dev(a["amp_r"], np.float32),
         dev(a["amp_i"], np.float32), dev(a["cx"], np.float64), dev(a["cy"], np.float64),
         dev(a["cz"], np.float64)]
    out_r = torch.zeros(npix, dtype=torch.float32, device=device)
    out_i = torch.zeros(npix, dtype=torch.float32, device=device)
    nb = cffi_stub.workspace_bytes(lib, n, a["Ny"], a["Nz"])
    ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=device)
    p = [x.data_ptr() for x in t]
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(repeats):
        cffi_stub.compute_scattering_gpu(
            lib, n, p[0], p[1], p[2], f["remove_forward"], p[3], p[4], p[5], p[6], a["Ny"],
            a["Nz"], p[7], p[8], p[9], a["k_val"], f["apply_pol"], f["pol_deg"], *f["pol_e"],
            f["apply_decay"], out_r.data_ptr(), out_i.data_ptr(), ws.data_ptr(), nb, stream)
    torch.cuda.synchronize()
    return (out_r.cpu().numpy() + 1j * out_i.cpu().numpy()).astype(np.complex128)


def oracle80(a, f, **extra):
    return O.kin_sum_c(**a, **f, precision="f80", **extra)


@pytest.fixture(scope="module")
def stub_lib():
    return cffi_stub.bind(_lib.LIB_PATH)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("fi", range(5))
@pytest.mark.parametrize("pi", range(2))
def test_golden_through_cffi_stub(device, stub_lib, name, fi, pi):
    a, f = golden_args(name), flags(fi, pi)
    print(f"[golden cffi] {name} f{fi} p{pi}", flush=True)
    got = stub_call(stub_lib, a, f, device)
    want = GOLD[f"{name}/out_f{fi}_p{pi}"].astype(np.complex128)
    err_gold = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    err80 = rel(got, oracle80(a, f))
    print(f"[golden cffi] {name} f{fi} p{pi}: gold {err_gold:.2e} f80 {err80:.2e}", flush=True)
    assert err_gold <= 5e-5 and err80 <= TOL


def test_cffi_stub_accumulates_and_rejects_small_workspace(device, stub_lib):
    a, f = golden_args("random64"), flags(4, 1)
    once = stub_call(stub_lib, a, f, device)
    twice = stub_call(stub_lib, a, f, device, repeats=2)
    assert rel(twice, 2 * once) <= 1e-6
    # a workspace smaller than the query: an error status, nothing launched
    t = torch.zeros(64 * 11, dtype=torch.float64, device=device)
    o = torch.zeros(64, dtype=torch.float32, device=device)
    p, q = t.data_ptr(), o.data_ptr()
    with pytest.raises(MemoryError):
        cffi_stub.compute_scattering_gpu(stub_lib, 64, p, p, p, 0, p, p, p, p, 8, 8, p, p, p,
                                         a["k_val"], 0, 0.0, 0.0, 1.0, 0.0, 0, q, q, p, 0)
    torch.cuda.synchronize()
    assert not torch.any(o != 0)


def test_cffi_large_sample_many_slices(device, stub_lib):
    """A sample split over many atom slices (and a ragged last LDS step)
    against the long-double oracle."""
    w = W.fcc_au_bragg111(n=9, npix=20, seed=4)
    a = W.oracle_inputs(w)
    f = dict(remove_forward=True, apply_pol=True, apply_decay=True, pol_deg=a.pop("pol_deg"),
             pol_e=a.pop("pol_e"))
    assert a["pos_m"].shape[0] % 256 != 0
    got = stub_call(stub_lib, a, f, device)
    e = rel(got, oracle80(a, f))
    print(f"[cffi fcc9 20px] rel {e:.2e}", flush=True)
    assert e <= TOL


def _staged_field(pos_A, codes, labels_consts, amp, coords_m, Ny, Nz, k, opts, device,
                  rot=None, trans=None):
    d_pos = torch.from_numpy(np.ascontiguousarray(pos_A, dtype=np.float32)).to(device)
    d_cod = torch.from_numpy(np.ascontiguousarray(codes, dtype=np.uint8)).to(device)
    d_amp = torch.from_numpy(np.ascontiguousarray(amp, dtype=np.complex64)).to(device)
    counts = np.bincount(codes, minlength=labels_consts["wk"].shape[0]).astype(np.int64)
    staged = K.stage(d_pos, d_cod, d_amp, counts, k, 1, Ny, Nz)
    coords = torch.from_numpy(np.ascontiguousarray(coords_m, dtype=np.float64)).to(device)
    out = torch.zeros((1, Nz, Ny), dtype=torch.complex128, device=device)
    rot = np.eye(3, dtype=np.float32)[None] if rot is None else rot
    trans = np.zeros((1, 3), np.float32) if trans is None else trans
    K.kin_sum(staged, K.species_table(labels_consts), opts, rot, trans, coords, out)
    torch.cuda.synchronize()
    return out[0].cpu().numpy().ravel(), staged.path_used


def _species_from_rows(a):
    rows = np.concatenate([a["f0p"], a["f0z"][:, None], a["sanr"][:, None],
                           a["sani"][:, None]], axis=1)
    uniq, codes = np.unique(rows, axis=0, return_inverse=True)
    consts = dict(wk=uniq[:, :11].astype(np.float32), f0z=uniq[:, 11].astype(np.float32),
                  fp=uniq[:, 12].astype(np.float32), fpp=uniq[:, 13].astype(np.float32))
    return consts, codes.astype(np.uint8).ravel()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("fi", range(5))
def test_golden_through_staged_sum(device, name, fi):
    """The staged drop-in path on the golden cases, automatic kernel choice;
    near30nm (pixels 30 nm from atoms 12 A apart) takes the exact kernel."""
    for pi in range(2):
        a, f = golden_args(name), flags(fi, pi)
        consts, codes = _species_from_rows(a)
        pos_A = (a["pos_m"] * 1e10).astype(np.float32)
        amp = (a["amp_r"] + 1j * a["amp_i"]).astype(np.complex64)
        coords = np.stack([a["cx"], a["cy"], a["cz"]])
        opts = K.kin_opts(a["k_val"], f["remove_forward"], f["apply_pol"],
                          (f["pol_deg"],) + f["pol_e"], f["apply_decay"])
        got, path = _staged_field(pos_A, codes, consts, amp, coords, a["Ny"], a["Nz"],
                                  a["k_val"], opts, device)
        want = GOLD[f"{name}/out_f{fi}_p{pi}"].astype(np.complex128)
        a80 = dict(a, pos_m=pos_A.astype(np.float64) / 1e10)
        err_gold = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
        err80 = rel(got, oracle80(a80, f))
        print(f"[golden staged] {name} f{fi} p{pi} path {path}: gold {err_gold:.2e} "
              f"f80 {err80:.2e}", flush=True)
        if name == "near30nm":
            assert path == 50
        assert err_gold <= 5e-5 and err80 <= TOL


@pytest.mark.parametrize("dist_a,pitch_a", [(2.0e6, 1.5e2), (500.0, 40.0)])
@pytest.mark.parametrize("fl", [dict(remove_forward=False, apply_pol=True, decay=False),
                                dict(remove_forward=True, apply_pol=True, decay=True)])
def test_near_field_crystal(device, dist_a, pitch_a, fl):
    """A 6^3-cell Au crystal (864 atoms, 24 A across) seen from 0.2 mm and
    from 50 nm: the far-field kernels decline, the automatic choice runs the exact
    kernel, which matches the long-double oracle."""
    w = W.fcc_au_bragg111(n=6, npix=16, seed=2)
    d = w["detector"]
    d.create_detector((16, 16), (pitch_a, pitch_a))
    d.position_detector_absolute(dist_a, np.rad2deg(2 * w["theta_b"]), 0.0)
    b, s, st = w["beam"], w["sample"], w["stage"]
    rot = np.asarray(st.rotation, dtype=np.float32)[None]
    tr = np.asarray(st.translation, dtype=np.float32)[None]
    A = b._kin_field(s, d.pixel_coordinates, d.shape, st, (rot, tr), fl["remove_forward"],
                     fl["apply_pol"], fl["decay"])[0].cpu().numpy().ravel()
    assert b._last_path == 50
    a = W.oracle_inputs(w)
    kw = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx",
                            "cy", "cz", "k_val", "Ny", "Nz", "pol_deg", "pol_e")}
    R = O.kin_sum_c(**kw, remove_forward=fl["remove_forward"], apply_pol=fl["apply_pol"],
                    apply_decay=fl["decay"], precision="f80")
    e = rel(A, R)
    print(f"[near field] {dist_a:.0e} A {fl}: rel {e:.2e}", flush=True)
    assert e <= TOL


def test_exact_forced_in_far_field(device):
    """Forced exact kernel on a rotated, translated far-field pose with three
    species (Beam.py's general kernel in exact mode)."""
    w = W.perovskite_screw_bragg200(n=6, npix=24, blocks=(2, 1, 1))
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    st.set_single_motor_value_absolute("phi", 0.3, degrees=True)
    b._kin_path = "exact"
    rot = np.asarray(st.rotation, dtype=np.float32)[None]
    tr = (np.asarray(st.translation, dtype=np.float32) + np.float32([3.0, -2.0, 1.0]))[None]
    A = b._kin_field(s, d.pixel_coordinates, d.shape, st, (rot, tr), True, True,
                     True)[0].cpu().numpy().ravel()
    assert b._last_path == 50
    a = W.oracle_inputs(w, frame_rot=rot[0], frame_trans=tr[0])
    kw = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx",
                            "cy", "cz", "k_val", "Ny", "Nz", "pol_deg", "pol_e")}
    R = O.kin_sum_c(**kw, remove_forward=True, apply_pol=True, apply_decay=True,
                    precision="f80")
    e = rel(A, R)
    print(f"[exact forced far field] rel {e:.2e}", flush=True)
    assert e <= TOL


def _one_atom_case():
    """One Au atom at the origin, 24^2 pixels of 75 um at 1 m around 2theta 30 deg."""
    from mosaic_amd import tables
    from mosaic_amd.detector import detector
    d = detector()
    d.create_detector((24, 24), (7.5e5, 7.5e5))
    d.position_detector_absolute(1e10, 30.0, 0.0)
    coords = np.asarray(d.pixel_coordinates, dtype=np.float64) / 1e10
    consts = tables.species_constants(["Au"], 10000.0)
    k = 2 * np.pi / tables.wavelength_from_energy_ev(10000.0)
    centre = coords[:, 12 * 24 + 12]
    u = coords / np.linalg.norm(coords, axis=0)
    c = centre / np.linalg.norm(centre)
    delta = np.arctan2(np.linalg.norm(np.cross(c[None, :], u.T), axis=1), u.T @ c)
    return coords, consts, k, centre, delta


@pytest.mark.parametrize("kind", [1, 2])
def test_analyser_one_atom_closed_form(device, kind):
    """For one atom at the origin every ray is its pixel's direction, so the
    analyser multiplies the plain field by its weight at the pixel's angle
    delta to the axis: the top-hat mask delta <= accept (3.5 pixels, between
    pixel rings), the Darwin factor 1 / (1 + (delta / hw)^2)."""
    coords, consts, k, centre, delta = _one_atom_case()
    pos = np.zeros((1, 3), np.float32)
    codes = np.zeros(1, np.uint8)
    amp = np.ones(1, np.complex64)
    plain_opts = K.kin_opts(k, apply_pol=True, pol=(1.0, 0.0, 1.0, 0.0), path="exact")
    plain, _ = _staged_field(pos, codes, consts, amp, coords, 24, 24, k, plain_opts, device)
    accept, hw = 3.5 * 7.5e-5, 1.5e-4
    opts = K.kin_opts(k, apply_pol=True, pol=(1.0, 0.0, 1.0, 0.0), analyser=kind,
                      analyser_axis=centre, accept_rad=accept, darwin_hw_rad=hw)
    got, path = _staged_field(pos, codes, consts, amp, coords, 24, 24, k, opts, device)
    assert path == 50 + kind
    wgt = (delta <= accept).astype(float) if kind == 1 else 1.0 / (1.0 + (delta / hw) ** 2)
    if kind == 1:
        assert 0 < np.count_nonzero(wgt) < wgt.size
    want = plain * wgt
    e = rel(got, want)
    print(f"[analyser one atom] kind {kind}: rel {e:.2e}", flush=True)
    assert e <= 1e-6


@pytest.mark.parametrize("kind", [1, 2])
def test_analyser_crystal_vs_oracle(device, kind):
    w = W.fcc_au_bragg111(n=6, npix=24, seed=5)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    coords = np.asarray(d.pixel_coordinates, dtype=np.float64)
    centre = coords[:, 12 * 24 + 12] / 1e10
    accept, hw = 3.5 * 7.5e-5, 1.5e-4
    mode = "top_hat" if kind == 1 else "darwin"
    A = b.interact_beam_gpu(s, coords, d.shape, st, apply_polarization=True,
                            spherical_decay=True, analyser_mode=mode,
                            analyser_acceptance_angle_rad=accept,
                            analyser_darwin_halfwidth_rad=hw).ravel()
    assert b._last_path == 50 + kind
    a = W.oracle_inputs(w)
    R = O.kin_sum_numpy(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                        a["amp_i"], a["cx"], a["cy"], a["cz"], a["k_val"], False, a["Ny"],
                        a["Nz"], True, a["pol_deg"], a["pol_e"], True,
                        dtype=np.longdouble,
                        analyser=(kind, centre, accept, hw)).astype(np.complex128)
    e = rel(A, R)
    print(f"[analyser crystal] kind {kind}: rel {e:.2e}", flush=True)
    assert e <= TOL
