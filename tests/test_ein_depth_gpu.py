"""GPU parity of the depth-dependent entrance field (mosaic_ein_depth_*)
against its oracle (oracle/ein.py), and the drop-in with use_depth_ein.

Bars: the column map A within 1e-5 relative (float32 atomics reorder the
column sums), the per-atom Ein within 2e-6 of the oracle's sampler on the same
maps, the kinematic field within 1e-5 of the long-double oracle fed the
oracle's Ein; the Ein cache returns the computed values bit for bit.
"""
import os

import numpy as np
import pytest
import torch

from mosaic_amd import kinematic as K
from mosaic_amd import tables
from mosaic_amd import workloads as W
from oracle import ein as OE
from oracle import kinematic as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(seed=0, n=20000):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-60, 60, (n, 3)).astype(np.float32)
    codes = rng.integers(0, 3, n).astype(np.uint8)
    consts = tables.species_constants(["Au", "Ti", "O"], 10000.0)
    fwd = np.stack([(consts["f0z"] + consts["fp"]).astype(np.float32), consts["fpp"]], axis=1)
    ang = 0.4
    R = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]],
                 np.float32)
    T = np.array([5.0, -3.0, 2.0], np.float32)
    return pos, codes, fwd, R, T


@pytest.mark.parametrize("radius", [0, 2])
def test_column_map_and_sampler_vs_oracle(device, radius):
    pos, codes, fwd, R, T = _setup()
    NyB, NzB, du, dv = 64, 48, 2.5, 3.0
    uc, vc = (NyB - 1) / 2.0, (NzB - 1) / 2.0
    lam = 1.2398419843320025
    d_pos = torch.from_numpy(pos).to(device)
    d_cod = torch.from_numpy(codes).to(device)
    de = K.DepthEin(NyB, NzB, device)
    s_min, s_max = de.bounds(d_pos, R, T)
    lab = OE.lab_positions(pos, R, T)
    assert (s_min, s_max) == OE.depth_bounds(lab)
    A = torch.zeros((NyB, NzB), dtype=torch.complex64, device=device)
    de.column_map(d_pos, d_cod, fwd, R, T, du, dv, uc, vc, lam, radius, A_out=A)
    sre, sim = OE.column_sums(lab, fwd[codes, 0], fwd[codes, 1], NyB, NzB, du, dv, uc, vc)
    A_ref = OE.column_map(sre, sim, du, dv, lam, radius)
    e_A = rel(A.cpu().numpy(), A_ref)
    rng = np.random.default_rng(1)
    E0 = (rng.random((NyB, NzB)) + 1j * rng.random((NyB, NzB))).astype(np.complex64)
    ein = de.sample(d_pos, R, T, torch.from_numpy(E0).to(device), du, dv, uc, vc, s_min,
                    s_max).cpu().numpy()
    tau, phi = OE.sampled_maps(A.cpu().numpy())
    ein_ref = OE.ein_depth(lab, tau, phi, E0, du, dv, uc, vc, s_min, s_max)
    e_E = rel(ein, ein_ref)
    print(f"[depth ein] radius {radius}: A rel {e_A:.2e}, Ein rel {e_E:.2e}", flush=True)
    assert np.any(ein == 0) and np.count_nonzero(ein) > 1000
    assert e_A <= 1e-5 and e_E <= 2e-6


def test_dropin_depth_ein_and_cache(device, tmp_path):
    w = W.fcc_au_bragg111(n=10, npix=16, seed=3, n_chunks=4)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    b.create_beam(10000.0, beam_size=(60.0, 60.0), beam_samples=(48, 40), pol_perp_rate=1.0)
    st.set_single_motor_value_absolute("phi", 3.0)
    kw = dict(apply_polarization=True, spherical_decay=False, use_depth_ein=True)
    A = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, **kw)
    # the oracle: depth window, map and Ein from oracle/ein.py, then the sum
    a = W.oracle_inputs(w)
    pos = np.concatenate([s.load_chunk_positions(c) for c in range(1, s.chunk_total + 1)])
    R = np.asarray(st.rotation, np.float32)
    T = np.asarray(st.translation, np.float32)
    lab = OE.lab_positions(pos, R, T)
    s_min, s_max = OE.depth_bounds(lab)
    consts = tables.species_constants(["Au"], b.energy)
    fr = np.full(pos.shape[0], np.float32(consts["f0z"][0] + consts["fp"][0]))
    fi = np.full(pos.shape[0], consts["fpp"][0])
    sre, sim = OE.column_sums(lab, fr, fi, b._beam_Ny, b._beam_Nz, b._beam_du, b._beam_dv,
                              b._beam_uc, b._beam_vc)
    Amap = OE.column_map(sre, sim, b._beam_du, b._beam_dv, b.wavelength * 1e10)
    tau, phi = OE.sampled_maps(Amap)
    amp = OE.ein_depth(lab, tau, phi, b._beam_E0_map, b._beam_du, b._beam_dv, b._beam_uc,
                       b._beam_vc, s_min, s_max)
    assert np.any(amp == 0) and np.std(np.abs(amp[amp != 0])) > 0
    a = W.oracle_inputs(w, amp=amp)
    args = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx",
                              "cy", "cz", "k_val", "Ny", "Nz", "pol_deg", "pol_e")}
    Rf = O.kin_sum_c(**args, remove_forward=False, apply_pol=True, apply_decay=False,
                     precision="f80").reshape(a["Nz"], a["Ny"])
    e = rel(A, Rf)
    plain = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, apply_polarization=True,
                                spherical_decay=False)
    print(f"[depth ein] drop-in vs oracle rel {e:.2e}; depth Ein changes the field by "
          f"{rel(plain, A):.2e}", flush=True)
    assert e <= 1e-5 and rel(plain, A) > 1e-4
    # the Ein cache: computed and written, then read back bit for bit
    cache = str(tmp_path / "ein_cache")
    A1 = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, ein_cache_dir=cache, **kw)
    assert b._last_ein_source == "computed"
    files = sorted(os.listdir(cache))
    assert len(files) == s.chunk_total and all(f.startswith("ein_chunk_") for f in files)
    A2 = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, ein_cache_dir=cache, **kw)
    assert b._last_ein_source == "cache" and np.array_equal(A1, A2)
    b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, ein_cache_dir=cache,
                        recompute_cache=True, **kw)
    assert b._last_ein_source == "computed"
    st.set_single_motor_value_absolute("phi", 3.5)          # a new pose: new key, old pruned
    b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, ein_cache_dir=cache, **kw)
    assert len(os.listdir(cache)) == s.chunk_total and sorted(os.listdir(cache)) != files
    st.set_single_motor_value_absolute("phi", 0.0)
