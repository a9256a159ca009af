"""Beam: the drop-in for the reference's kinematic scattering and propagation calls.

Public methods keep the reference's names, argument order, defaults and
return conventions:

* ``create_beam`` (Beam.py:152-243), ``_polarization_state`` (117-149);
* ``atomic_scattering_kinematic`` (Beam.py:6409-6496) -> (Nz, Ny) complex64;
* ``interact_beam_gpu`` (Beam.py:6028-6407) -> (Nz, Ny) complex64;
* ``atomic_direct_interaction`` (Beam.py:7771-7824) -> writes the detector;
* ``wavefield_propagation`` (Beam.py:8024-8095) -> writes the detector or
  returns the field.

Everything runs on the GPU through ``libmosaic_amd.so``.  Where the
reference would fall back to its CPU path (``use_gpu=False``, no device)
this build raises: the CPU kernels are not part of it.

Deliberate differences, each bounded by the tests:
* the sample-to-lab transform ``pos @ R.T + T`` is applied exactly (FP64)
  inside the kernels instead of rounding lab positions to float32 first;
* chunk fields are accumulated in complex128 and cast once at the end;
* form factor, polarization and 1/r decay are evaluated once per
  (256-atom tile, pixel) instead of per atom -- exact to first order in
  tile radius / distance, with a decline beyond 1e-6.
"""
import numpy as np
import torch

from . import kinematic as K
from . import tables

_ANALYSER_OFF = ("off", "none", "disabled")


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("mosaic_amd runs on the GPU only: no HIP device is visible")


class beam:
    def __init__(self, directory=None):
        self.directory = directory
        self._direction = None
        self._energy = None
        self._wavelength = None
        self._pol_perp_rate = 0.5
        self._pol_direction = None
        self._pol_degree = None
        self._h = tables.H_PLANCK
        self._c = tables.C_LIGHT
        self._q = tables.Q_E
        self._hq = self._h / self._q
        self._workspace = None
        # kinematic kernel selection: "auto", "tile" or "rows" (tests pin it)
        self._kin_path = "auto"

    # -------------------------------------------------------------- tables
    parse_f0_db_all = staticmethod(lambda database_name="f0_WaasKirf.dat": tables.parse_f0_db_all())
    parse_f1f2_db_all = staticmethod(
        lambda database_name="f1f2_CromerLiberman.dat": tables.parse_f1f2_db_all())
    get_f1f2_from_params = staticmethod(tables.get_f1f2_from_params)
    _build_f0_zero_dict = staticmethod(tables.build_f0_zero_dict)

    # --------------------------------------------------------------- setup
    def _set_polarization(self, pol_perp_rate, pol_direction=None, pol_degree=None):
        self._pol_perp_rate = float(np.clip(pol_perp_rate, 0.0, 1.0))
        self._pol_direction = (None if pol_direction is None
                               else np.asarray(pol_direction, dtype=np.float64).reshape(3))
        self._pol_degree = None if pol_degree is None else float(np.clip(pol_degree, 0.0, 1.0))

    def _polarization_state(self):
        khat = np.asarray(self._direction if self._direction is not None else (1.0, 0.0, 0.0),
                          dtype=np.float64)
        khat = khat / np.linalg.norm(khat)
        p_rate, ex, ey, ez = tables.polarization_from_rate(self._pol_perp_rate)
        p = p_rate if self._pol_degree is None else self._pol_degree
        e = np.array([ex, ey, ez]) if self._pol_direction is None else self._pol_direction
        e = np.asarray(e, dtype=np.float64)
        e = e - np.dot(e, khat) * khat
        if np.linalg.norm(e) < 1e-12:
            for axis in (np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])):
                e = axis - np.dot(axis, khat) * khat
                if np.linalg.norm(e) >= 1e-12:
                    break
        e = e / np.linalg.norm(e)
        return float(p), float(e[0]), float(e[1]), float(e[2])

    def create_beam(self, energy, eV=True, beam_shape="rectangular", beam_size=(1000.0, 1000.0),
                    beam_samples=(256, 256), beam_profile="uniform", gaussian_waist=None,
                    pol_perp_rate=0.5, pol_direction=None, pol_degree=None):
        self._direction = np.array([1.0, 0.0, 0.0], dtype=np.float32)
        if not eV:
            energy = energy / self._q
        self._energy = float(energy)
        self._wavelength = self._hq * self._c / self._energy
        k = 2.0 * np.pi / self._wavelength
        self._kx_scalar = np.float32(k)
        self._beam_shape = str(beam_shape).lower()
        self._beam_size = (float(beam_size[0]), float(beam_size[1]))
        self._beam_samples = (int(beam_samples[0]), int(beam_samples[1]))
        self._beam_profile = str(beam_profile).lower()
        self._gauss_waist = gaussian_waist
        self._set_polarization(pol_perp_rate, pol_direction, pol_degree)
        self._beam_e1 = np.array([0.0, 1.0, 0.0], dtype=np.float32)
        self._beam_e2 = np.array([0.0, 0.0, 1.0], dtype=np.float32)
        self._init_beam_grid()

    def _init_beam_grid(self):
        """Transverse grid and entrance field E0 (Beam.py:245-264)."""
        Ny, Nz = (max(1, int(v)) for v in self._beam_samples)
        Sy, Sz = self._beam_size
        self._beam_Ny, self._beam_Nz = Ny, Nz
        self._beam_du = float(Sy) / Ny
        self._beam_dv = float(Sz) / Nz
        self._beam_uc = (Ny - 1) * 0.5
        self._beam_vc = (Nz - 1) * 0.5
        u = (np.arange(Ny, dtype=np.float32) - self._beam_uc) * self._beam_du
        v = (np.arange(Nz, dtype=np.float32) - self._beam_vc) * self._beam_dv
        U, V = np.meshgrid(u, v, indexing="ij")
        if self._beam_shape == "circular":
            mask = ((U / max(0.5 * Sy, 1e-9)) ** 2 + (V / max(0.5 * Sz, 1e-9)) ** 2) <= 1.0
        else:
            mask = np.ones_like(U, dtype=bool)
        if self._beam_profile == "gaussian":
            wy, wz = (self._gauss_waist if self._gauss_waist is not None else (0.5 * Sy, 0.5 * Sz))
            wy = max(float(wy), 1e-6)
            wz = max(float(wz), 1e-6)
            A0 = np.exp(-((U / wy) ** 2 + (V / wz) ** 2)).astype(np.float32) * mask.astype(np.float32)
        else:
            A0 = mask.astype(np.float32)
        self._beam_u_centers = u
        self._beam_v_centers = v
        self._beam_E0_map = (A0.astype(np.float32) + 0.0j).astype(np.complex64)

    def set_wavefield(self, wavefield):
        arr = np.asarray(wavefield)
        if arr.shape != (self._beam_Ny, self._beam_Nz):
            raise ValueError(f"set_wavefield: expected shape {(self._beam_Ny, self._beam_Nz)}, "
                             f"got {arr.shape}")
        arr = np.ascontiguousarray(arr, dtype=np.complex64)
        if not np.all(np.isfinite(arr.real)) or not np.all(np.isfinite(arr.imag)):
            raise ValueError("set_wavefield: wavefield contains non-finite values (NaN/Inf)")
        self._beam_E0_map = arr

    @property
    def wavelength(self):
        return self._wavelength

    @property
    def energy(self):
        return self._energy

    # ------------------------------------------------------------ helpers
    def _species_setup(self, sample):
        """Global label table and per-chunk code remaps."""
        labels = []
        index = {}
        remaps = []
        for cid in range(1, int(sample.chunk_total or 0) + 1):
            codes, lab = sample.load_chunk_species(cid, use_gpu=False, codes=True)
            rm = np.zeros(max(1, len(lab)), dtype=np.uint8)
            for i, el in enumerate(lab):
                el = str(el)
                if el not in index:
                    index[el] = len(labels)
                    labels.append(el)
                rm[i] = index[el]
            remaps.append(rm)
        if len(labels) > 16:
            raise ValueError("at most 16 species per sample are supported")
        return labels, remaps

    def _kin_field(self, sample, coords_A, shape, stage, frames, remove_forward,
                   apply_polarization, spherical_decay, out_dtype=torch.complex128):
        """Sum every chunk of ``sample`` for the poses ``frames`` = (rot, trans).

        Chunks are uploaded from pinned host buffers on a copy stream,
        overlapping the previous chunk's sum.  Returns a (F, Nz, Ny) device
        tensor.
        """
        _require_gpu()
        if self._wavelength is None:
            raise ValueError("Beam wavelength is not set; call create_beam first")
        dev = torch.device("cuda", torch.cuda.current_device())
        Ny, Nz = int(shape[0]), int(shape[1])
        rot, trans = frames
        F = rot.shape[0]
        k_val = float(2.0 * np.pi / self._wavelength)
        coords = torch.from_numpy(
            np.ascontiguousarray(np.asarray(coords_A, dtype=np.float64) / 1e10)).to(dev)
        out = torch.zeros((F, Nz, Ny), dtype=out_dtype, device=dev)
        n_chunks = int(sample.chunk_total or 0)
        if n_chunks == 0:
            return out
        labels, remaps = self._species_setup(sample)
        consts = tables.species_constants(labels, self._energy)
        spc_tab = K.species_table(consts)
        opts = K.kin_opts(k_val, remove_forward, apply_polarization, self._polarization_state(),
                          spherical_decay, path=self._kin_path)
        E0 = torch.from_numpy(self._beam_E0_map).to(dev)
        compute = torch.cuda.current_stream()
        copy = torch.cuda.Stream()
        staged_prev = None

        def upload(cid):
            pos = np.ascontiguousarray(sample.load_chunk_positions(cid, use_gpu=False),
                                       dtype=np.float32)
            codes, _ = sample.load_chunk_species(cid, use_gpu=False, codes=True)
            codes = remaps[cid - 1][np.asarray(codes, dtype=np.int64)] if len(codes) else \
                np.zeros(0, dtype=np.uint8)
            counts = np.bincount(codes, minlength=len(labels)).astype(np.int64)
            h_pos = torch.from_numpy(pos).pin_memory()
            h_cod = torch.from_numpy(np.ascontiguousarray(codes, dtype=np.uint8)).pin_memory()
            with torch.cuda.stream(copy):
                d_pos = h_pos.to(dev, non_blocking=True)
                d_cod = h_cod.to(dev, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(copy)
            return d_pos, d_cod, counts, ev, (h_pos, h_cod)

        pending = upload(1)
        for cid in range(1, n_chunks + 1):
            d_pos, d_cod, counts, ev, hold = pending
            compute.wait_event(ev)
            d_pos.record_stream(compute)
            d_cod.record_stream(compute)
            if cid < n_chunks:
                pending = upload(cid + 1)
            n = int(d_pos.shape[0])
            if n == 0:
                continue
            amp = K.ein_bilinear(d_pos, rot[0], trans[0], E0, self._beam_du, self._beam_dv,
                                 self._beam_uc, self._beam_vc)
            staged = K.stage(d_pos, d_cod, amp, counts, k_val, F, Ny, Nz,
                             workspace=self._workspace)
            if staged_prev is not None and staged.workspace is staged_prev.workspace:
                # pixel tables sit at the head of the workspace and survive staging
                staged.pixels_key = staged_prev.pixels_key
                staged.pixel_geom = staged_prev.pixel_geom
            self._workspace = staged.workspace
            K.kin_sum(staged, spc_tab, opts, rot, trans, coords, out)
            self._last_path = staged.path_used
            self._last_stats = (staged.radius_wl, tuple(staged.pixel_geom))
            staged_prev = staged
            del hold
        del staged_prev
        return out

    # ------------------------------------------------------- public calls
    def interact_beam_gpu(self, sample, measurement_positions, measurement_shape, stage,
                          remove_forward=False, use_depth_ein=False, ein_cache_dir=None,
                          recompute_cache=False, apply_polarization=False, spherical_decay=True,
                          analyser_mode="off", analyser_acceptance_angle_rad=0.0,
                          analyser_darwin_halfwidth_rad=0.0):
        if use_depth_ein:
            raise NotImplementedError("depth-dependent Ein is not part of this build")
        off = (analyser_mode is None
               or (isinstance(analyser_mode, str) and analyser_mode.strip().lower() in _ANALYSER_OFF)
               or (not isinstance(analyser_mode, str) and analyser_mode == 0))
        if not off:
            raise NotImplementedError("analyser acceptance is not part of this build")
        Ny, Nz = measurement_shape
        rot = np.asarray(stage.rotation, dtype=np.float32).reshape(1, 3, 3)
        trans = np.asarray(stage.translation, dtype=np.float32).reshape(1, 3)
        mp = measurement_positions
        if isinstance(mp, torch.Tensor):
            mp = mp.detach().cpu().numpy()
        out = self._kin_field(sample, mp, (Ny, Nz), stage, (rot, trans), remove_forward,
                              apply_polarization, spherical_decay)
        return out[0].to(torch.complex64).cpu().numpy()

    def atomic_scattering_kinematic(self, sample, detector, stage, offset=None, use_gpu=True,
                                    remove_forward=False, use_depth_ein=False,
                                    ein_cache_dir=None, recompute_cache=False,
                                    apply_polarization=False, spherical_decay=False,
                                    analyser_mode="off", analyser_acceptance_angle_rad=0.0,
                                    analyser_darwin_halfwidth_rad=0.0):
        if not use_gpu:
            raise NotImplementedError("the CPU scattering path is not part of this build")
        Ny, Nz = detector.shape
        field = self.interact_beam_gpu(
            sample, detector.pixel_coordinates, (Ny, Nz), stage, remove_forward=remove_forward,
            use_depth_ein=use_depth_ein, ein_cache_dir=ein_cache_dir,
            recompute_cache=recompute_cache, apply_polarization=apply_polarization,
            spherical_decay=spherical_decay, analyser_mode=analyser_mode,
            analyser_acceptance_angle_rad=analyser_acceptance_angle_rad,
            analyser_darwin_halfwidth_rad=analyser_darwin_halfwidth_rad)
        return (field - offset) if offset is not None else field

    def atomic_direct_interaction(self, sample, detector, stage, scattering=True,
                                  sc_kwargs=None, transmission=False, tr_kwargs=None,
                                  use_gpu=True):
        if transmission:
            raise NotImplementedError("the transmission path is not part of this build")
        Ny, Nz = detector.shape
        field = np.zeros((Nz, Ny), dtype=np.complex64)
        if scattering:
            kw = dict(sc_kwargs or {})
            kw.setdefault("use_gpu", use_gpu)
            field = field + self.atomic_scattering_kinematic(sample, detector, stage, **kw)
        detector.input_pixel_values(field.astype(np.complex64))

    def kinematic_scan(self, sample, detector, stage, motor, angles_deg, remove_forward=False,
                       apply_polarization=False, spherical_decay=False, out_dtype=torch.complex64):
        """All poses of a rocking scan in one batched launch.

        ``motor`` is a stage rotation motor; each angle (degrees) is set
        absolutely and the stage rotation/translation recorded, as one
        ``Experiment.scan_nD`` step would (Experiment.py:285-320).  Returns a
        (n_angles, Nz, Ny) device tensor.  The incident amplitude is sampled
        at the first pose.
        """
        start = stage.get_single_motor_value(motor, degrees=True)
        rots, trs = [], []
        try:
            for a in angles_deg:
                stage.set_single_motor_value_absolute(motor, float(a), degrees=True)
                rots.append(np.asarray(stage.rotation, dtype=np.float32))
                trs.append(np.asarray(stage.translation, dtype=np.float32))
        finally:
            stage.set_single_motor_value_absolute(motor, start, degrees=True)
        rot = np.stack(rots)
        trans = np.stack(trs)
        out = self._kin_field(sample, detector.pixel_coordinates, detector.shape, stage,
                              (rot, trans), remove_forward, apply_polarization, spherical_decay,
                              out_dtype=torch.complex128)
        return out.to(out_dtype)

    # --------------------------------------------------------- propagation
    def free_space_propagator(self, step_max=None, pad_factor=1.0, padding_mode="edge",
                              pad_constant=0.0):
        """Callback ``(F, dy, dz, z, lens=None) -> F`` for optics.apply_stack.

        F is a (Nz, Ny) complex64 device tensor; the reference's closure
        (Beam.py:8062-8072) with the automatic alias-free sub-stepping.
        """
        from . import propagation as P
        wl = self._wavelength

        def _propagate_free_space(F, dym, dzm, z, lens=None):
            if not isinstance(F, torch.Tensor):
                F = torch.as_tensor(np.ascontiguousarray(F, dtype=np.complex64),
                                    device=torch.device("cuda", torch.cuda.current_device()))
            return P.propagate(F.to(torch.complex64).contiguous(), dym, dzm, wl, z,
                               step_max=step_max, pad_factor=pad_factor,
                               padding_mode=padding_mode, pad_constant=pad_constant, lens=lens)

        _propagate_free_space.accepts_lens = True
        return _propagate_free_space

    def wavefield_propagation(self, detector, optics, use_gpu=True, step_max=None,
                              pad_factor=1.0, padding_mode="edge", pad_constant=0.0,
                              save_field=True):
        """Propagate the detector field through an optics stack (Beam.py:8024-8095)."""
        if not use_gpu:
            raise NotImplementedError("the CPU propagation path is not part of this build")
        _require_gpu()
        dy, dz = np.asarray(detector.pixel_size, dtype=np.float64) * 1e-10
        E = detector.pixel_values
        if E is None:
            raise ValueError("detector.pixel_values is None. Run atomic_direct_interaction() "
                             "first to populate the detector field before calling "
                             "wavefield_propagation().")
        if self._wavelength is None:
            raise ValueError("Beam wavelength is not set. Call beam.create_beam() first.")
        prop = self.free_space_propagator(step_max, pad_factor, padding_mode, pad_constant)
        E = optics.apply_stack(field=E, dx=dy, dy=dz, wavelength=self._wavelength,
                               propagate_free_space=prop, use_gpu=True)
        if save_field is True:
            detector.input_pixel_values(E.astype(np.complex64))
        else:
            return E.astype(np.complex64)

