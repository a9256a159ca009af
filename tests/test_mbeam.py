"""The M-channel incident amplitude amp = sum_m E_m e^{2 pi i g_m . r}
(Beam.py:2283-2309, r the lab position in metres, g in cycles per metre).

Oracle pins (CPU): g = 0 gives sum_m E_m exactly, the reference's
bit-identity of M = 1, g = 0 with the single-amplitude path; a point whose
g . r is a quarter cycle turns E by +i (a sign slip gives -i); atoms on the
lattice g . r in Z give sum_m E_m whatever their distance from the origin; a
phase built with r in Angstrom instead of metres fails the quarter-cycle pin.
GPU: mosaic_mbeam_amplitude against the oracle within 2e-6 relative (the
hardware sine/cosine, 1e-6 absolute), the pose applied as rot @ pos + trans;
the amplitudes then feed the general sum (mosaic_scatter), which matches the
long-double oracle fed the oracle's amplitudes within 1e-5 relative.
"""
import numpy as np
import pytest
import torch

from oracle import ein as OE


def _quarter_case():
    g = np.array([[2.5e9, 0.0, 0.0]])            # cycles per metre
    r = np.array([[1e-10, 0.0, 0.0]])            # g . r = 0.25
    return r, g


def test_mbeam_oracle_pins():
    rng = np.random.default_rng(0)
    E = (rng.standard_normal((7, 4)) + 1j * rng.standard_normal((7, 4)))
    r = rng.uniform(-1e-7, 1e-7, (7, 3))
    assert np.array_equal(OE.mbeam_amplitude(r, E, np.zeros((4, 3))), E.sum(axis=1))
    rq, gq = _quarter_case()
    assert np.allclose(OE.mbeam_amplitude(rq, [[1.0 + 0j]], gq), [1j], atol=1e-15)
    # lattice points of g = (1/a, 0, 0) with a = 4 A, up to 2.5 um from the origin
    a = 4e-10
    n = np.array([0, 1, -3, 1000, 6250, -6249])
    rl = np.column_stack([n * a, rng.uniform(-1e-6, 1e-6, n.size), np.zeros(n.size)])
    G = np.array([[1.0 / a, 0.0, 0.0], [0.0, 0.0, 3.0 / a]])
    El = np.ones((n.size, 2), complex) * np.array([1.0 + 2j, -0.5 + 0j])
    assert np.allclose(OE.mbeam_amplitude(rl, El, G), 0.5 + 2j, atol=1e-12)


def test_mbeam_checks_catch_sign_and_units():
    rq, gq = _quarter_case()
    wrong_sign = np.conj(OE.mbeam_amplitude(rq, [[1.0 + 0j]], gq))
    wrong_units = OE.mbeam_amplitude(rq * 1e10, [[1.0 + 0j]], gq)
    assert not np.allclose(wrong_sign, [1j], atol=1e-6)
    assert not np.allclose(wrong_units, [1j], atol=1e-6)


@pytest.mark.gpu
def test_mbeam_amplitude_gpu(device):
    from mosaic_amd import kinematic as K
    rng = np.random.default_rng(3)
    n, M = 50000, 5
    pos = rng.uniform(-400, 400, (n, 3)).astype(np.float32)
    ang = 0.3
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]],
                 np.float32)
    T = np.array([12.0, -7.5, 3.25], np.float32)
    E = (rng.standard_normal((n, M)) + 1j * rng.standard_normal((n, M))).astype(np.complex64)
    a = 4.08e-10
    g = np.array([[0, 0, 0], [1 / a, 0, 0], [0, 2 / a, 0], [1 / a, 1 / a, 1 / a],
                  [-2 / a, 0, 4 / a]], np.float64)
    out = K.mbeam_amplitude(torch.from_numpy(pos).to(device), R, T,
                            torch.from_numpy(E).to(device), g)
    torch.cuda.synchronize()
    lab = (pos.astype(np.float64) @ R.astype(np.float64).T + T.astype(np.float64)) * 1e-10
    ref = OE.mbeam_amplitude(lab, E, g)
    e = float(np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref))
    print(f"[mbeam] N={n} M={M}: rel {e:.2e}", flush=True)
    assert e <= 2e-6
    # g = 0 everywhere: exactly the channel sum in float64 rounded once
    out0 = K.mbeam_amplitude(torch.from_numpy(pos).to(device), R, T,
                             torch.from_numpy(E).to(device), np.zeros((M, 3)))
    assert np.array_equal(out0.cpu().numpy(), E.astype(np.complex128).sum(axis=1)
                          .astype(np.complex64))


@pytest.mark.gpu
def test_mbeam_amplitudes_feed_the_general_sum(device):
    """The general sum (mosaic_scatter) fed the GPU's M-channel amplitudes
    against the long-double oracle fed the oracle's: relative L2 <= 1e-5."""
    from mosaic_amd import kinematic as K
    from mosaic_amd import workloads as W
    from oracle import kinematic as O
    w = W.fcc_au_bragg111(n=6, npix=16, seed=4)
    st = w["stage"]
    st.set_single_motor_value_absolute("phi", 0.02)
    a0 = W.oracle_inputs(w)
    n = a0["pos_m"].shape[0]
    rng = np.random.default_rng(5)
    M = 3
    E = (rng.standard_normal((n, M)) + 1j * rng.standard_normal((n, M))).astype(np.complex64)
    a = 4.078e-10
    g = np.array([[0, 0, 0], [1 / a, 1 / a, 1 / a], [-1 / a, 1 / a, 1 / a]], np.float64)
    R = np.asarray(st.rotation, np.float32)
    T = np.asarray(st.translation, np.float32)
    s = np.concatenate([w["sample"].load_chunk_positions(c)
                        for c in range(1, w["sample"].chunk_total + 1)]).astype(np.float32)
    amp_gpu = K.mbeam_amplitude(torch.from_numpy(s).to(device), R, T,
                                torch.from_numpy(E).to(device), g)
    amp_ref = OE.mbeam_amplitude(a0["pos_m"], E, g)
    a = W.oracle_inputs(w, amp=amp_ref)
    dev = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(device)
    out_r = torch.zeros(a["Ny"] * a["Nz"], dtype=torch.float32, device=device)
    out_i = torch.zeros_like(out_r)
    K.scatter(dev(a["pos_m"], np.float64), dev(a["f0p"], np.float32),
              dev(a["f0z"], np.float32), dev(a["sanr"], np.float32),
              dev(a["sani"], np.float32), torch.real(amp_gpu).contiguous(),
              torch.imag(amp_gpu).contiguous(), dev(a["cx"], np.float64),
              dev(a["cy"], np.float64), dev(a["cz"], np.float64), a["k_val"], False,
              a["Ny"], a["Nz"], True, a["pol_deg"], a["pol_e"], False, out_r, out_i)
    torch.cuda.synchronize()
    got = out_r.cpu().numpy().astype(np.float64) + 1j * out_i.cpu().numpy()
    args = {k: a[k] for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx",
                              "cy", "cz", "k_val", "Ny", "Nz", "pol_deg", "pol_e")}
    ref = O.kin_sum_c(**args, remove_forward=False, apply_pol=True, apply_decay=False,
                      precision="f80")
    e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"[mbeam] general sum fed M=3 amplitudes: rel {e:.2e}", flush=True)
    assert e <= 1e-5
