"""Chunk files in the reference's layout and the thermal modes on load.

* write an in-memory sample as ``atomic_positions_<cid>.npy``,
  ``atomic_species_<cid>.npy`` (codes) and ``sample_metadata.json``, read it
  back with ``sample.from_directory``: identical chunks, labels and counts;
  species stored as labels (older files) read as the same codes;
* Einstein mode: sigma^2 = hbar / (2 m omega) coth(hbar omega / 2 k_B T) has
  the classical limit k_B T / (m omega^2) at T >> theta_E and the zero-point
  value hbar / (2 m omega) at T = 0; per-species masses pick each atom's own;
* Debye mode: phi(x) = (3 / x^4) int_0^x t^3 / (e^t - 1) dt tends to 1 / x as
  x -> 0 (classical <u^2> = 9 hbar^2 T / (m k_B theta_D^2)) and to
  pi^4 / (5 x^4) for large x;
* every mode draws from the (seed, chunk) stream: a chunk loads the same
  displacements every time, with the sigma the model gives.
"""
import json
import os

import numpy as np

from mosaic_amd import sample as S


def _sample():
    rng = np.random.default_rng(0)
    pos = [rng.uniform(-50, 50, (n, 3)).astype(np.float32) for n in (300, 0, 257, 1000)]
    codes = [rng.integers(0, 3, p.shape[0]).astype(np.uint8) for p in pos]
    return S.sample(pos, codes, ["Sr", "Ti", "O"])


def test_chunk_files_roundtrip(tmp_path):
    s = _sample()
    s.write_chunks(str(tmp_path))
    assert sorted(os.listdir(tmp_path))[:3] == ["atomic_positions_1.npy", "atomic_positions_2.npy",
                                                "atomic_positions_3.npy"]
    meta = json.load(open(tmp_path / "sample_metadata.json"))
    assert meta["chunk_total"] == 4 and meta["species_labels"] == ["Sr", "Ti", "O"]
    d = S.sample.from_directory(str(tmp_path))
    assert d.chunk_total == 4 and d.atom_count == s.atom_count
    for cid in range(1, 5):
        assert d.chunk_atom_count(cid) == s.chunk_atom_count(cid)
        assert np.array_equal(d.load_chunk_positions(cid), s.load_chunk_positions(cid))
        c1, l1 = d.load_chunk_species(cid, codes=True)
        c0, l0 = s.load_chunk_species(cid, codes=True)
        assert c1.dtype == np.uint8 and np.array_equal(c1, c0) and list(l1) == list(l0)
        assert np.array_equal(d.load_chunk_species(cid), s.load_chunk_species(cid))


def test_species_stored_as_labels(tmp_path):
    s = _sample()
    s.write_chunks(str(tmp_path))
    labels = s.load_chunk_species(4)
    np.save(tmp_path / "atomic_species_4.npy", labels)           # an older file: labels
    d = S.sample.from_directory(str(tmp_path))
    codes, lab = d.load_chunk_species(4, codes=True)
    assert np.array_equal(lab[codes], labels)


def test_einstein_limits():
    s = _sample()
    s.set_temperature_einstein(0.0, mass_amu=196.97, theta_E_K=130.0)
    m = 196.97 * S.AMU_KG
    w = S.K_B * 130.0 / S.HBAR
    zp = s.thermal_sigma("einstein", 0.0, None, 1)[0]
    assert np.isclose(zp, np.sqrt(S.HBAR / (2 * m * w)) / 1e-10, rtol=1e-12)
    hot = s.thermal_sigma("einstein", 1.3e5, None, 1)[0]
    assert np.isclose(hot, np.sqrt(S.K_B * 1.3e5 / (m * w * w)) / 1e-10, rtol=1e-6)
    # per species: Sr, Ti, O masses give each atom its own sigma
    s.set_temperature_einstein(300.0, species_mass_amu={"Sr": 87.62, "Ti": 47.87, "O": 16.0},
                               species_theta_E_K={"Sr": 200.0, "Ti": 300.0, "O": 500.0})
    sig = s.thermal_sigma("einstein", 300.0, 3, s.chunk_atom_count(3))
    lab = s.load_chunk_species(3)
    for el, mm, th in (("Sr", 87.62, 200.0), ("O", 16.0, 500.0)):
        ref = s.__class__()
        ref.set_temperature_einstein(300.0, mass_amu=mm, theta_E_K=th)
        assert np.allclose(sig[lab == el], ref.thermal_sigma("einstein", 300.0, None, 1)[0],
                           rtol=1e-12)


def test_debye_limits():
    x = 1e-3
    assert np.isclose(S.sample.debye_phi(x), 1.0 / x, rtol=1e-3)
    assert np.isclose(S.sample.debye_phi(60.0), np.pi ** 4 / (5 * 60.0 ** 4), rtol=1e-12)
    assert np.isclose(S.sample.debye_phi(40.0), np.pi ** 4 / (5 * 40.0 ** 4), rtol=1e-9)
    s = _sample()
    s.set_temperature_debye(3.0e4, mass_amu=63.55, theta_D_K=343.0)
    sig = s.thermal_sigma("debye", 3.0e4, None, 1)[0]
    m = 63.55 * S.AMU_KG
    classical = np.sqrt(9 * S.HBAR ** 2 * 3.0e4 / (m * S.K_B * 343.0 ** 2)) / 1e-10
    assert np.isclose(sig, classical, rtol=1e-3)


def test_thermal_draws_keyed_and_scaled():
    s = _sample()
    for setup in (lambda: s.set_temperature("gaussian", 0.1, None, 7),
                  lambda: s.set_temperature_einstein(300.0, mass_amu=50.0, theta_E_K=250.0,
                                                     seed=7),
                  lambda: s.set_temperature_debye(300.0, mass_amu=50.0, theta_D_K=250.0,
                                                  seed=7)):
        setup()
        a = s.load_chunk_positions(4)
        b = s.load_chunk_positions(4)
        assert np.array_equal(a, b) and a.dtype == np.float32
        raw = s.load_chunk_positions(4, raw=True)
        d = (a - raw).astype(np.float64)
        sig = s.thermal_sigma(s.temp_params[0], s.temp_params[1], 4, raw.shape[0])[0]
        assert abs(d.std() / sig - 1.0) < 0.05
        c = s.load_chunk_positions(1)
        assert not np.allclose(c - s.load_chunk_positions(1, raw=True), d[:300])


def test_chunk_positions_view(tmp_path):
    """The copy-free view the drop-in gathers from equals the loaded positions,
    in memory and from chunk files, and is withheld when thermal
    displacements apply on load."""
    s = _sample()
    s.write_chunks(str(tmp_path))
    d = S.sample.from_directory(str(tmp_path))
    for smp in (s, d):
        for cid in range(1, smp.chunk_total + 1):
            v = smp.chunk_positions_view(cid)
            assert v is not None and np.array_equal(np.asarray(v), smp.load_chunk_positions(cid))
    s.set_temperature("gaussian", 0.05, None, 3)
    assert s.chunk_positions_view(1) is None
