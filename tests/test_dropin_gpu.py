"""The public drop-in calls against the oracle, with the reference's defaults.

* ``interact_beam_gpu`` (Beam.py:6028-6043): spherical_decay=True and
  apply_polarization=False by default; returns (Nz, Ny) complex64;
* ``atomic_scattering_kinematic`` (Beam.py:6409-6496): decay and polarization
  off by default, ``offset`` subtracted;
* ``atomic_direct_interaction`` (Beam.py:7771-7824): writes the detector's
  field, forwarding sc_kwargs;
* thermal displacements drawn on chunk load enter the sum;
* a scan through a Gaussian beam narrower than the sample samples the incident
  amplitude at every pose (the poses then run one at a time), equal to single
  poses.
Bar: relative L2 <= 1e-5 against the long-double oracle (complex64 outputs
add ~6e-8).
"""
import numpy as np
import pytest
import torch

from mosaic_amd import workloads as W
from oracle import kinematic as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle(w, remove_forward, apply_pol, decay, **kw):
    a = W.oracle_inputs(w, **kw)
    args = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx",
                              "cy", "cz", "k_val", "Ny", "Nz", "pol_deg", "pol_e")}
    R = O.kin_sum_c(**args, remove_forward=remove_forward, apply_pol=apply_pol,
                    apply_decay=decay, precision="f80")
    return R.reshape(a["Nz"], a["Ny"])


def test_interact_beam_gpu_defaults(device):
    w = W.perovskite_screw_bragg200(n=8, npix=24, blocks=(2, 2, 2))
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    A = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st)
    assert A.shape == (d.shape[1], d.shape[0]) and A.dtype == np.complex64
    e = rel(A, oracle(w, False, False, True))
    print(f"[dropin] interact_beam_gpu defaults (decay on): rel {e:.2e}", flush=True)
    assert e <= TOL


def test_atomic_scattering_kinematic_offset(device):
    w = W.fcc_au_bragg111(n=8, npix=20, seed=6)          # Gaussian thermal draw on load
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    rng = np.random.default_rng(0)
    offset = (rng.standard_normal(d.shape[::-1]) * 1e-12).astype(np.complex64)
    A = b.atomic_scattering_kinematic(s, d, st, offset=offset)
    R = oracle(w, False, False, False) - offset
    e = rel(A, R)
    print(f"[dropin] atomic_scattering_kinematic with offset: rel {e:.2e}", flush=True)
    assert e <= TOL
    with pytest.raises(NotImplementedError):
        b.atomic_scattering_kinematic(s, d, st, use_gpu=False)


def test_atomic_direct_interaction_writes_detector(device):
    w = W.fcc_au_bragg111(n=7, npix=16, seed=2, n_chunks=3)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    st.set_single_motor_value_absolute("phi", 0.01)
    b.atomic_direct_interaction(s, d, st, sc_kwargs=dict(apply_polarization=True,
                                                         remove_forward=True,
                                                         spherical_decay=True))
    A = np.asarray(d.pixel_values)
    e = rel(A, oracle(w, True, True, True))
    print(f"[dropin] atomic_direct_interaction: rel {e:.2e}", flush=True)
    assert e <= TOL


def test_scan_resamples_incident_amplitude_per_pose(device):
    """A 12 A Gaussian beam on a 6-cell (24 A) crystal: rotating the sample
    moves atoms through the profile, so each pose has its own amplitudes."""
    w = W.fcc_au_bragg111(n=6, npix=16, seed=1, sigma=0.0)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    b.create_beam(10000.0, beam_size=(30.0, 30.0), beam_samples=(64, 64),
                  beam_profile="gaussian", gaussian_waist=(12.0, 12.0), pol_perp_rate=1.0)
    angles = [-20.0, 0.0, 25.0]
    frames = b.kinematic_scan(s, d, st, "phi", angles, apply_polarization=True,
                              out_dtype=torch.complex128).cpu().numpy()
    for i, a in enumerate(angles):
        st.set_single_motor_value_absolute("phi", a)
        single = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st,
                                     apply_polarization=True, spherical_decay=False)
        e = rel(frames[i], single)
        print(f"[dropin] gaussian beam scan pose {a}: rel vs single pose {e:.2e}", flush=True)
        assert e <= 1e-6
    st.set_single_motor_value_absolute("phi", 0.0)
    assert rel(frames[0], frames[2]) > 1e-2


def test_chunk_files_full_perovskite(device, tmp_path):
    """The 5M-atom perovskite written as 64 chunk files (the reference's
    layout) and read back through the drop-in: the same field as the
    in-memory sample, bit for bit, and the same field as the oracle on 32
    sampled pixels."""
    import os
    import time
    from mosaic_amd.sample import sample
    w = W.perovskite_screw_bragg200(n=100, npix=512)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    s.write_chunks(str(tmp_path))
    assert len([f for f in os.listdir(tmp_path) if f.startswith("atomic_positions_")]) == 64
    files = sample.from_directory(str(tmp_path))
    A_mem = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, apply_polarization=True,
                                spherical_decay=False)
    b.interact_beam_gpu(files, d.pixel_coordinates, d.shape, st, apply_polarization=True,
                        spherical_decay=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A_file = b.interact_beam_gpu(files, d.pixel_coordinates, d.shape, st,
                                 apply_polarization=True, spherical_decay=False)
    dt = time.perf_counter() - t0
    print(f"[dropin] 64 chunk files, 5M atoms x 512^2: {dt * 1e3:.1f} ms per call", flush=True)
    assert np.array_equal(A_mem, A_file)
    a = W.oracle_inputs(w)
    sel = np.random.default_rng(2).choice(a["Ny"] * a["Nz"], size=32, replace=False)
    R = O.kin_sum_c(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                    a["amp_i"], a["cx"][sel], a["cy"][sel], a["cz"][sel], a["k_val"], False, 32,
                    1, True, a["pol_deg"], a["pol_e"], False, precision="f64",
                    threads=min(16, os.cpu_count() or 1))
    e = rel(A_file.ravel()[sel], R)
    print(f"[dropin] 64 chunk files: 32 pixels vs FP64 oracle rel {e:.2e}", flush=True)
    assert e <= TOL


def test_pinned_sample_same_field(device):
    """A sample moved into page-locked memory (sample.pin_memory(): chunks
    copied by DMA, no host copy) gives the same field bit for bit; with a
    thermal draw on load its chunks take the host path again."""
    w = W.perovskite_screw_bragg200(n=12, npix=32)
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    A0 = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, spherical_decay=False)
    s.pin_memory()
    assert s.pinned_chunk(1) is not None
    A1 = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, spherical_decay=False)
    assert np.array_equal(A0, A1)
    s.set_temperature("gaussian", 0.05, None, 3)
    assert s.pinned_chunk(1) is None
    A2 = b.interact_beam_gpu(s, d.pixel_coordinates, d.shape, st, spherical_decay=False)
    assert not np.array_equal(A2, A1)
    w2 = W.fcc_au_bragg111(n=6, npix=16, seed=0, sigma=0.0)
    b2, s2, d2, st2 = w2["beam"], w2["sample"], w2["detector"], w2["stage"]
    B0 = b2.interact_beam_gpu(s2, d2.pixel_coordinates, d2.shape, st2)
    s2.pin_memory()
    B1 = b2.interact_beam_gpu(s2, d2.pixel_coordinates, d2.shape, st2)
    assert np.array_equal(B0, B1)
