"""The synthetic workloads the benchmarks and parity tests run.

Each builder returns a dict with ``sample``, ``detector``, ``stage``,
``beam`` (and extras), built from the documented recipe:

* ``fcc_au_bragg111`` -- FCC Au (a = 4.08 A) n^3 cells centred on the
  origin, [111] along the (111) scattering vector, Gaussian thermal
  displacements sigma = 0.08 A drawn on load from ``seed``; 10 keV sigma-
  polarized beam; square detector of ``npix`` 75 um pixels at 1.0 m
  centred on 2theta(111).  n = 20, npix = 128 is the 32,000-atom
  CPU-runnable case.
* ``perovskite_screw_bragg200`` -- SrTiO3-type basis (a = 3.905 A), n^3
  cells in (4x4x4) spatial chunks, screw dislocation u_z = b atan2(y, x)
  / 2 pi (b = 4 A) along the crystal z axis through the centre; (200)
  reflection; npix^2 detector at 1 m.  n = 100, npix = 512 is the
  5,000,000-atom precision workload.
* ``fcc_fault_scan111`` -- FCC Au (nx, ny, nz) cells with one intrinsic
  stacking fault on a (111) plane at mid-height, between atomic planes (atoms above shift by
  a/6 [1 1 -2]); ``n_angles`` phi rotations uniformly in +-0.05 deg around the
  (111) Bragg condition; npix^2 detector.  (50, 50, 100) cells = 1,000,000
  atoms, 64 angles, 256^2 pixels is the batched rocking-scan workload.
"""
import numpy as np

from . import sample as S
from . import tables
from .beam import beam as Beam
from .detector import detector as Detector
from .stage import stage as Stage

ENERGY_EV = 10000.0
PIXEL_A = 7.5e5          # 75 um
DISTANCE_A = 1.0e10      # 1.0 m


def _detector(npix, two_theta_rad, pixel_a=PIXEL_A, distance_a=DISTANCE_A):
    d = Detector()
    d.create_detector((npix, npix), (pixel_a, pixel_a))
    d.position_detector_absolute(distance_a, two_theta_rad, 0.0, degrees=False)
    return d


def _beam(pol_perp_rate=1.0):
    b = Beam()
    b.create_beam(ENERGY_EV, pol_perp_rate=pol_perp_rate)
    return b


def fcc_au_bragg111(n=20, npix=128, seed=0, sigma=0.08, n_chunks=1):
    a = 4.08
    lam = tables.wavelength_from_energy_ev(ENERGY_EV)
    th = S.bragg_angle(a / np.sqrt(3.0), lam)
    pos, _ = S.lattice(a, (n, n, n), S.FCC_BASIS)
    R = S.orient_for_reflection((1, 1, 1), th)
    pos = pos @ R.T
    codes = np.zeros(pos.shape[0], dtype=np.uint8)
    cp, cc = S.split_chunks(pos.astype(np.float32), codes, n_chunks)
    smp = S.sample(cp, cc, ["Au"])
    if sigma > 0:
        smp.set_temperature("gaussian", sigma, 1.0, seed)
    return dict(sample=smp, detector=_detector(npix, 2 * th), stage=Stage(), beam=_beam(1.0),
                theta_b=th, name=f"fcc_au_{n}c_{npix}px")


def perovskite_screw_bragg200(n=100, npix=512, blocks=(4, 4, 4), b=4.0, seed=0):
    a = 3.905
    lam = tables.wavelength_from_energy_ev(ENERGY_EV)
    th = S.bragg_angle(a / 2.0, lam)
    pos, bidx = S.lattice(a, (n, n, n), S.PEROVSKITE_BASIS)
    # species: 0 Sr, 1 Ti, 2 O (basis order Sr, Ti, O, O, O)
    codes = np.array([0, 1, 2, 2, 2], dtype=np.uint8)[bidx]
    pos[:, 2] += S.screw_dislocation_uz(pos, b=b)
    R = S.orient_for_reflection((1, 0, 0), th)
    cp, cc = S.block_chunks(pos, codes, a, (n, n, n), blocks)
    cp = [(p @ R.T).astype(np.float32) for p in cp]
    smp = S.sample(cp, cc, ["Sr", "Ti", "O"])
    return dict(sample=smp, detector=_detector(npix, 2 * th), stage=Stage(), beam=_beam(1.0),
                theta_b=th, name=f"perovskite_screw_{n}c_{npix}px")


def fcc_fault_scan111(cells=(50, 50, 100), npix=256, n_angles=64, half_range_deg=0.05,
                      sigma=0.0, seed=0, fault=True):
    a = 4.08
    lam = tables.wavelength_from_energy_ev(ENERGY_EV)
    th = S.bragg_angle(a / np.sqrt(3.0), lam)
    pos, _ = S.lattice(a, cells, S.FCC_BASIS)
    if fault:
        # plane a quarter (111) spacing off the centre: between atomic planes
        pos = S.stacking_fault_shift(pos, a, plane_offset=0.25 * a / np.sqrt(3.0))
    R = S.orient_for_reflection((1, 1, 1), th)
    pos = (pos @ R.T).astype(np.float32)
    smp = S.sample([pos], [np.zeros(pos.shape[0], dtype=np.uint8)], ["Au"])
    if sigma > 0:
        smp.set_temperature("gaussian", sigma, 1.0, seed)
    angles = np.linspace(-half_range_deg, half_range_deg, n_angles)
    return dict(sample=smp, detector=_detector(npix, 2 * th), stage=Stage(), beam=_beam(1.0),
                theta_b=th, angles_deg=angles, motor="phi",
                name=f"fcc_fault_{cells[0]}x{cells[1]}x{cells[2]}_{npix}px_{n_angles}ang")


def oracle_inputs(w, frame_rot=None, frame_trans=None, amp=None, atom_subset=None):
    """Flat arguments for the oracle's kinematic sum of workload ``w``.

    Lab positions are R @ s + T evaluated in float64 from the float32 sample
    positions and float32 stage, the convention of the HIP path.
    """
    smp, det, stg, bm = w["sample"], w["detector"], w["stage"], w["beam"]
    R = np.asarray(stg.rotation if frame_rot is None else frame_rot, dtype=np.float32)
    T = np.asarray(stg.translation if frame_trans is None else frame_trans, dtype=np.float32)
    pos_l, cod_l, labels_all = [], [], []
    index = {}
    for cid in range(1, smp.chunk_total + 1):
        p = smp.load_chunk_positions(cid)
        c, lab = smp.load_chunk_species(cid, codes=True)
        rm = []
        for el in lab:
            el = str(el)
            if el not in index:
                index[el] = len(labels_all)
                labels_all.append(el)
            rm.append(index[el])
        pos_l.append(p)
        cod_l.append(np.asarray(rm, dtype=np.int64)[np.asarray(c, dtype=np.int64)])
    s = np.concatenate(pos_l).astype(np.float64)
    codes = np.concatenate(cod_l)
    if atom_subset is not None:
        s = s[atom_subset]
        codes = codes[atom_subset]
    lab_m = (s @ R.astype(np.float64).T + T.astype(np.float64)) / 1e10
    consts = tables.species_constants(labels_all, bm.energy)
    n = s.shape[0]
    if amp is None:
        amp = np.ones(n, dtype=np.complex64)
    pc = np.asarray(det.pixel_coordinates, dtype=np.float64) / 1e10
    Ny, Nz = det.shape
    return dict(pos_m=lab_m, f0p=consts["wk"][codes], f0z=consts["f0z"][codes],
                sanr=consts["fp"][codes], sani=consts["fpp"][codes],
                amp_r=np.real(amp).astype(np.float32), amp_i=np.imag(amp).astype(np.float32),
                cx=pc[0], cy=pc[1], cz=pc[2], k_val=2.0 * np.pi / bm.wavelength,
                Ny=Ny, Nz=Nz, pol_deg=bm._polarization_state()[0],
                pol_e=bm._polarization_state()[1:])


def gaussian_bump_field(pos, amplitude=0.2, width=200.0, rng=None, centre=None, direction=None):
    """Per-atom displacement u = amplitude exp(-|r - r0|^2 / (2 width^2)) e_hat (float32).

    r0 and e_hat are drawn from ``rng`` when not given: r0 uniformly in the
    middle half of the sample's bounding box, e_hat uniformly on the sphere.
    """
    from .sample import gaussian_bump
    rng = np.random.default_rng(0) if rng is None else rng
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    if centre is None:
        centre = lo + (hi - lo) * (0.25 + 0.5 * rng.random(3))
    if direction is None:
        direction = rng.standard_normal(3)
    return gaussian_bump(pos.astype(np.float64), centre, width, amplitude,
                         direction).astype(np.float32)


def fcc_strain_fit(cells=(63, 63, 63), npix=256, n_angles=16, half_range_deg=0.05, seed=0):
    """The displacement-fitting workload: FCC Au, (111) rocking poses, a smooth
    Gaussian strain bump (0.2 A, width 200 A) as the current field and a second
    seeded bump as the target."""
    w = fcc_fault_scan111(cells=cells, npix=npix, n_angles=n_angles,
                          half_range_deg=half_range_deg, fault=False)
    pos = w["sample"].load_chunk_positions(1)
    w["u"] = gaussian_bump_field(pos, 0.2, 200.0, np.random.default_rng(seed))
    w["u_target"] = gaussian_bump_field(pos, 0.2, 200.0, np.random.default_rng(seed + 1))
    w["name"] = f"fcc_strain_fit_{cells[0]}x{cells[1]}x{cells[2]}_{npix}px_{n_angles}ang"
    return w
