"""Parity at the full perovskite workload: 5,000,000 atoms (Pb, Ti, O) with a
screw dislocation, 512 x 512 pixels, Bragg 200 geometry.

The oracle cannot evaluate all 1.3e12 pairs in test time, so the full-size
path (nine chunked global k-d levels, units of tiles, the degree-6 moment sum)
is held to two checks that do not depend on the size:
- superposition: the field of all atoms equals the sum of the fields of the
  even- and odd-indexed atoms, each half staged on its own (its own k-d tree,
  tiles and units, so the expansion errors of the three sums are unrelated);
- sampled pixels: 32 random pixels of the full image against the FP64 C
  oracle summed over all 5,000,000 atoms.
The sampled pixels meet the 1e-5 relative bar of the small-size tests. The
superposition is held to 1e-6: the three sums share the rounding of the pixel
tables and phases, which cancels in the difference (1.1e-7 measured), while a
lost or doubled 256-atom tile moves the Bragg peak by about 5e-5.
"""
import os

import numpy as np
import pytest
import torch

from mosaic_amd import kinematic as K
from mosaic_amd import tables
from mosaic_amd import workloads as W
from oracle import kinematic as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL_SUPERPOSITION = 1e-6


def _gather(w):
    """All atoms of the in-memory sample with species codes in first-seen order."""
    smp = w["sample"]
    labels, index, pos, codes = [], {}, [], []
    for cid in range(1, smp.chunk_total + 1):
        p = smp.load_chunk_positions(cid)
        c, lab = smp.load_chunk_species(cid, codes=True)
        remap = []
        for el in lab:
            el = str(el)
            if el not in index:
                index[el] = len(labels)
                labels.append(el)
            remap.append(index[el])
        pos.append(p)
        codes.append(np.asarray(remap, dtype=np.uint8)[np.asarray(c, dtype=np.int64)])
    return np.concatenate(pos).astype(np.float32), np.concatenate(codes), labels


@pytest.fixture(scope="module")
def full(device):
    w = W.perovskite_screw_bragg200(n=100, npix=512)
    b, d, st = w["beam"], w["detector"], w["stage"]
    pos, codes, labels = _gather(w)
    assert pos.shape[0] == 5_000_000 and d.shape == (512, 512)
    k = 2.0 * np.pi / b.wavelength
    spc = K.species_table(tables.species_constants(labels, b.energy))
    opts = K.kin_opts(k, False, True, b._polarization_state(), False)
    rot = np.asarray(st.rotation, np.float32)[None]
    tr = np.asarray(st.translation, np.float32)[None]
    coords = torch.from_numpy(np.ascontiguousarray(d.pixel_coordinates / 1e10)).to(device)
    E0 = torch.from_numpy(b._beam_E0_map).to(device)
    Ny, Nz = d.shape

    def field(p, c):
        d_pos = torch.from_numpy(np.ascontiguousarray(p)).to(device)
        d_cod = torch.from_numpy(np.ascontiguousarray(c)).to(device)
        amp = K.ein_bilinear(d_pos, rot[0], tr[0], E0, b._beam_du, b._beam_dv, b._beam_uc,
                             b._beam_vc)
        counts = np.bincount(c, minlength=len(labels)).astype(np.int64)
        staged = K.stage(d_pos, d_cod, amp, counts, k, 1, Ny, Nz)
        out = torch.zeros((1, Nz, Ny), dtype=torch.complex128, device=device)
        K.kin_sum(staged, spc, opts, rot, tr, coords, out)
        torch.cuda.synchronize()
        return out[0].cpu().numpy().ravel(), staged.path_used, staged.radius_wl

    A, path, radius = field(pos, codes)
    print(f"[full] 5M atoms x 512^2: path {path}, largest tile radius {radius:.2f} wl",
          flush=True)
    return dict(w=w, pos=pos, codes=codes, A=A, path=path, field=field)


def test_fullsize_superposition(full):
    pos, codes = full["pos"], full["codes"]
    Ae, pe, _ = full["field"](pos[0::2], codes[0::2])
    Ao, po, _ = full["field"](pos[1::2], codes[1::2])
    A = full["A"]
    e = float(np.linalg.norm(A - (Ae + Ao)) / np.linalg.norm(A))
    print(f"[full] superposition (paths {full['path']}, {pe}, {po}): rel={e:.3e}", flush=True)
    assert full["path"] % 100 == 76          # the matrix-core moment sum
    assert e <= TOL_SUPERPOSITION


def test_fullsize_sampled_pixels_vs_fp64_oracle(full):
    a = W.oracle_inputs(full["w"])
    rng = np.random.default_rng(11)
    sel = rng.choice(a["Ny"] * a["Nz"], size=32, replace=False)
    R = O.kin_sum_c(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                    a["amp_i"], a["cx"][sel], a["cy"][sel], a["cz"][sel], a["k_val"], False,
                    32, 1, True, a["pol_deg"], a["pol_e"], False, precision="f64",
                    threads=min(16, os.cpu_count() or 1))
    e = float(np.linalg.norm(full["A"][sel] - R) / np.linalg.norm(R))
    print(f"[full] 32 sampled pixels vs FP64 oracle over 5M atoms: rel={e:.3e}", flush=True)
    assert e <= TOL
