"""Parity at the full rocking-scan and displacement-fitting workloads.

Rocking scan: 1,000,000-atom FCC Au (50 x 50 x 100 cells) with a (111)
stacking fault, 64 phi rotations in +-0.05 deg about the (111) Bragg angle,
256 x 256 pixels, all poses in one launch (beam.kinematic_scan):
- 3 frames x 32 random pixels against the FP64 C oracle summed over all
  atoms at each frame's pose (bar 1e-5, as the small-size tests);
- superposition: the scan of all atoms against the sum of the scans of the
  even- and odd-indexed atoms, each staged on its own (bar 1e-6; a lost
  256-atom tile moves the Bragg peak by ~2.5e-4).

Displacement fitting: 1,000,188-atom FCC Au (63^3 cells) with a Gaussian
strain bump (0.2 A, width 200 A), 16 poses x 256^2 pixels,
L = sum ((|A|^2 - I_t) / max I_t)^2 with I_t from a second seeded bump:
- the hand-written adjoint dL/du on 100 random atoms against the oracle's
  reference gradient (exact per-pair phase gradient, long-double atom terms,
  the GPU field's upstream gradient g), bar 1e-5;
- that reference against central finite differences of L (h = 1e-3 A) on the
  same atoms, formed from the atom's long-double terms (bar 1e-4).
"""
import os

import numpy as np
import pytest
import torch

from mosaic_amd import sample as S
from mosaic_amd import tables
from mosaic_amd import workloads as W
from mosaic_amd.fitting import KinematicModel
from oracle import adjoint as AO
from oracle import kinematic as O

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _poses(st, angles):
    rots, trs = [], []
    for a in angles:
        st.set_single_motor_value_absolute("phi", float(a))
        rots.append(np.asarray(st.rotation, np.float32))
        trs.append(np.asarray(st.translation, np.float32))
    st.set_single_motor_value_absolute("phi", 0.0)
    return np.stack(rots), np.stack(trs)


@pytest.fixture(scope="module")
def scan(device):
    w = W.fcc_fault_scan111(cells=(50, 50, 100), npix=256, n_angles=64)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    assert s.atom_count == 1_000_000 and d.shape == (256, 256)
    frames = b.kinematic_scan(s, d, st, "phi", w["angles_deg"], apply_polarization=True,
                              out_dtype=torch.complex128)
    torch.cuda.synchronize()
    print(f"[full scan] 1M atoms x 64 poses x 256^2: path {b._last_path}", flush=True)
    return dict(w=w, frames=frames.cpu().numpy().reshape(64, -1), path=b._last_path)


def test_scan_sampled_pixels_vs_fp64_oracle(scan):
    w = scan["w"]
    rots, trs = _poses(w["stage"], w["angles_deg"])
    rng = np.random.default_rng(21)
    errs = []
    for f in (0, 31, 63):
        a = W.oracle_inputs(w, frame_rot=rots[f], frame_trans=trs[f])
        sel = rng.choice(a["Ny"] * a["Nz"], size=32, replace=False)
        R = O.kin_sum_c(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                        a["amp_i"], a["cx"][sel], a["cy"][sel], a["cz"][sel], a["k_val"], False,
                        32, 1, True, a["pol_deg"], a["pol_e"], False, precision="f64",
                        threads=THREADS)
        errs.append(rel(scan["frames"][f][sel], R))
        print(f"[full scan] frame {f}: 32 pixels vs FP64 oracle rel={errs[-1]:.3e}", flush=True)
    assert scan["path"] % 100 == 76          # the matrix-core moment sum
    assert max(errs) <= 1e-5


def test_scan_superposition(scan):
    w = scan["w"]
    b, d, st = w["beam"], w["detector"], w["stage"]
    pos = w["sample"].load_chunk_positions(1)
    halves = []
    for part in (pos[0::2], pos[1::2]):
        smp = S.sample([part], [np.zeros(part.shape[0], np.uint8)], ["Au"])
        halves.append(b.kinematic_scan(smp, d, st, "phi", w["angles_deg"],
                                       apply_polarization=True,
                                       out_dtype=torch.complex128).cpu().numpy().reshape(64, -1))
    A = scan["frames"]
    e = rel(A, halves[0] + halves[1])
    print(f"[full scan] superposition over 64 frames: rel={e:.3e}", flush=True)
    assert e <= 1e-6


def test_fitting_gradient_vs_reference_and_fd(device):
    w = W.fcc_strain_fit(cells=(63, 63, 63), npix=256, n_angles=16)
    b, d, st = w["beam"], w["detector"], w["stage"]
    pos = w["sample"].load_chunk_positions(1)
    assert pos.shape[0] == 1_000_188
    rots, trs = _poses(st, w["angles_deg"])
    m = KinematicModel(pos, np.zeros(pos.shape[0], np.uint8), ["Au"], b, d, rots, trs)
    with torch.no_grad():
        It = (m.field(torch.from_numpy(w["u_target"]).to(device)).abs() ** 2)
    s = float(It.max())
    uu = torch.from_numpy(w["u"]).to(device).requires_grad_(True)
    A = m.field(uu)
    L = torch.sum(((A.abs() ** 2 - It) / s) ** 2)
    L.backward()
    torch.cuda.synchronize()
    print(f"[full adjoint] forward path {m.path_used}", flush=True)
    grad = uu.grad.cpu().numpy()
    A_np = A.detach().cpu().numpy().reshape(16, -1)
    It_np = It.cpu().numpy().reshape(16, -1)
    g = 4 * (np.abs(A_np) ** 2 - It_np) * A_np / s ** 2
    k = 2 * np.pi / b.wavelength
    pol = b._polarization_state()
    c = tables.species_constants(["Au"], b.energy)
    pc = d.pixel_coordinates / 1e10
    p_all = (pos + w["u"]).astype(np.float32)
    rng = np.random.default_rng(31)
    atoms = rng.choice(pos.shape[0], size=100, replace=False)
    args = (c["wk"][0], float(c["f0z"][0]), float(c["fp"][0]), float(c["fpp"][0]),
            np.complex64(1.0), rots, trs, pc, k, pol, THREADS)
    ref, fd, gpu = [], [], []
    h = 1e-3
    sq = np.sqrt(s)
    for j in atoms:
        cj = AO.atom_terms(p_all[j], *args)
        ref.append(AO.reference_gradient(cj, AO.phase_gradient(p_all[j], rots, trs, pc, k), g))
        row = []
        for ax in range(3):
            e = np.zeros(3)
            e[ax] = h
            cp = AO.atom_terms(p_all[j] + e, *args)
            cm = AO.atom_terms(p_all[j] - e, *args)
            row.append(AO.fd_gradient_intensity_loss(A_np / sq, It_np / s, cj / sq, cp / sq,
                                                     cm / sq, h))
        fd.append(row)
        gpu.append(grad[j])
    ref, fd, gpu = np.array(ref), np.array(fd), np.array(gpu)
    e_ref = rel(gpu, ref)
    e_fd = rel(ref, fd)
    print(f"[full adjoint] 100 atoms: GPU vs reference rel={e_ref:.3e}; reference vs FD "
          f"rel={e_fd:.3e}", flush=True)
    assert np.all(np.isfinite(grad))
    assert e_fd < 1e-4
    assert e_ref < 1e-5
