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
* all chunks are staged together and summed in one launch (all poses of a
  scan too), accumulated in complex128 and cast once at the end;
* the kernel is chosen by error bounds: the far-field forms hoist form
  factor, polarization and 1/r decay per (tile, pixel) or per pixel where
  that is exact to ~1e-7; elsewhere, and with an analyser, the exact
  per-pair kernel runs.
"""
import hashlib
from concurrent.futures import ThreadPoolExecutor
import json
import os

import numpy as np
import torch

from . import kinematic as K
from . import tables

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
        # kinematic kernel selection: "auto", "exact", "tile", "rows", "mma" or
        # "poly" (tests pin it)
        self._kin_path = "auto"
        # Gaussian blur radius [beam pixels] of the depth-Ein column map; the
        # reference's driver uses 0 (Beam.py:6240-6247)
        self._ein_kernel_radius = 0

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
    # Atoms staged together at most: the staging indexes atoms with 32 bits;
    # larger samples are summed in groups of whole chunks.
    MAX_GROUP_ATOMS = 1 << 30

    def _pinned(self, name, n, dtype):
        """A cached pinned host buffer of at least n elements."""
        bufs = self.__dict__.setdefault("_pin_bufs", {})
        b = bufs.get(name)
        if b is None or b.numel() < n or b.dtype != dtype:
            b = torch.empty(max(int(n), 1), dtype=dtype, pin_memory=True)
            bufs[name] = b
        return b

    def _gather_chunks(self, sample, chunk_ids, labels, index, dev, copy):
        """Load the chunks ``chunk_ids`` (positions with the sample's thermal
        displacements, species codes remapped onto the global label table)
        into one pinned host buffer and copy each chunk to the device on the
        ``copy`` stream as soon as it is loaded, so that loading the next
        chunks overlaps the transfer.  The label table is built first, in
        chunk order; the chunks are then loaded by a pool of threads (NumPy
        releases the GIL for the copies) straight into their slices of the
        pinned buffer.  Returns device positions (N, 3) float32, codes (N,)
        uint8, the per-species counts and the copies' event."""
        sizes = [int(sample.chunk_atom_count(c)) for c in chunk_ids]
        total = sum(sizes)
        ev_prev = getattr(self, "_pin_event", None)
        if ev_prev is not None:
            ev_prev.synchronize()        # the buffers' previous copies are done
        hp = self._pinned("pos", 3 * total, torch.float32)
        hc = self._pinned("codes", total, torch.uint8)
        d_pos = torch.empty((total, 3), dtype=torch.float32, device=dev)
        d_cod = torch.empty(total, dtype=torch.uint8, device=dev)
        copy.wait_stream(torch.cuda.current_stream(dev))
        hp_np, hc_np = hp.numpy(), hc.numpy()
        # species: each chunk's codes and its map onto the global labels
        offs, maps, code_arrs = [], [], []
        off = 0
        for cid, n in zip(chunk_ids, sizes):
            codes, lab = sample.load_chunk_species(cid, use_gpu=False, codes=True)
            rm = np.zeros(max(1, len(lab)), dtype=np.uint8)
            for i, el in enumerate(lab):
                el = str(el)
                if el not in index:
                    if len(labels) >= 16:
                        raise ValueError("at most 16 species per sample are supported")
                    index[el] = len(labels)
                    labels.append(el)
                rm[i] = index[el]
            offs.append(off)
            maps.append(None if np.array_equal(rm[:len(lab)], np.arange(len(lab))) else rm)
            code_arrs.append(codes)
            off += n
        view_of = getattr(sample, "chunk_positions_view", None)
        pinned_of = getattr(sample, "pinned_chunk", None)
        # pinned chunks whose codes are already global go by DMA, no host copy
        direct = {}
        if pinned_of is not None:
            for k, cid in enumerate(chunk_ids):
                if sizes[k] and maps[k] is None:
                    pc = pinned_of(cid)
                    if pc is not None:
                        direct[k] = pc

        def load(k):
            """Fill chunk k's slices of the pinned buffers; returns its
            per-species counts on the global label table."""
            cid, n, o = chunk_ids[k], sizes[k], offs[k]
            if n == 0:
                return None
            pos = view_of(cid) if view_of is not None else None
            if pos is None:
                pos = sample.load_chunk_positions(cid, use_gpu=False)
            np.copyto(hp_np[3 * o:3 * (o + n)].reshape(n, 3),
                      np.asarray(pos).reshape(n, 3), casting="same_kind")
            codes, rm = code_arrs[k], maps[k]
            dst = hc_np[o:o + n]
            if rm is None:                                      # codes already global
                np.copyto(dst, codes, casting="unsafe")
            else:                                               # codes checked on load
                np.take(rm, codes, out=dst, mode="clip")
            local = np.bincount(codes.reshape(-1), minlength=1)
            glob = np.zeros(16, dtype=np.int64)
            np.add.at(glob, np.arange(local.size) if rm is None else rm[:local.size], local)
            return glob

        workers = self.__dict__.get("_loader")
        if workers is None:
            workers = ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1)))
            self._loader = workers
        counts = np.zeros(16, dtype=np.int64)
        if direct:
            with torch.cuda.stream(copy):
                for k, (tp, tc, cnt) in direct.items():
                    o, n = offs[k], sizes[k]
                    d_pos[o:o + n].copy_(tp, non_blocking=True)
                    d_cod[o:o + n].copy_(tc, non_blocking=True)
                    counts[:len(cnt)] += cnt
        futures = [None if k in direct else workers.submit(load, k)
                   for k in range(len(chunk_ids))]
        # copies of runs of loaded chunks, about an eighth of the atoms each
        batch = max(1, total // 8)
        lo = 0
        hp2 = hp[:3 * total].view(total, 3) if total else None
        for k, fut in enumerate(futures):
            if fut is None:                     # copied by DMA above
                hi = offs[k]
                if hi > lo:                     # the host-loaded run before it
                    with torch.cuda.stream(copy):
                        d_pos[lo:hi].copy_(hp2[lo:hi], non_blocking=True)
                        d_cod[lo:hi].copy_(hc[lo:hi], non_blocking=True)
                lo = offs[k] + sizes[k]
                continue
            try:
                c = fut.result()
            except BaseException:
                # no worker may still write into the pinned buffers, which the
                # next call reuses
                for f in futures:
                    if f is not None:
                        f.cancel()
                for f in futures:
                    if f is not None and not f.cancelled():
                        try:
                            f.result()
                        except BaseException:
                            pass
                raise
            if c is not None:
                counts += c
            hi = offs[k] + sizes[k]
            if hi > lo and (hi - lo >= batch or k + 1 == len(futures)):
                with torch.cuda.stream(copy):
                    d_pos[lo:hi].copy_(hp2[lo:hi], non_blocking=True)
                    d_cod[lo:hi].copy_(hc[lo:hi], non_blocking=True)
                lo = hi
        ev = torch.cuda.Event()
        ev.record(copy)
        self._pin_event = ev
        counts = counts[:max(1, len(labels))].copy()
        return d_pos, d_cod, counts, ev

    # --------------------------------------------------- depth-dependent Ein
    def _ein_cache_key(self, sample, rot, trans, s_min, s_max, kernel_radius=0):
        """SHA1 naming one pose's cached depth-dependent Ein files: the beam,
        the pose, the depth window, the sample's chunk files (size and
        modification time) and thermal settings (the reference's key,
        Beam.py:5520-5564)."""
        files = []
        sdir = str(getattr(sample, "directory", "") or "")
        for cid in range(1, int(getattr(sample, "chunk_total", 0) or 0) + 1):
            p = os.path.join(sdir, f"atomic_positions_{cid}.npy")
            try:
                st = os.stat(p)
                files.append([cid, int(st.st_size), int(st.st_mtime_ns)])
            except OSError:
                files.append([cid, None, None])
        temp = getattr(sample, "temp_params", None)
        gw = self._gauss_waist
        key = dict(
            E_eV=float(self._energy), lam=float(self._wavelength),
            direction=[float(x) for x in self._direction],
            stage_R=np.asarray(rot, dtype=float).round(7).tolist(),
            stage_T=[float(x) for x in np.asarray(trans, dtype=float)],
            beam_size=[float(x) for x in self._beam_size],
            beam_samples=[int(self._beam_Ny), int(self._beam_Nz)],
            beam_profile=self._beam_profile,
            gauss_waist=[None if gw is None else float(gw[0]), None if gw is None else float(gw[1])],
            s_min=float(s_min), s_max=float(s_max),
            sample_dir=os.path.abspath(sdir) if sdir else "",
            chunk_files=files, enable_temp=bool(getattr(sample, "enable_temp", False)),
            temp_params=None if temp is None else [str(x) for x in temp],
            thermal_expansion=bool(getattr(sample, "_thermal_expansion_enabled", False)),
            kernel_radius=int(kernel_radius))
        return hashlib.sha1(json.dumps(key, sort_keys=True).encode("utf-8")).hexdigest()

    @staticmethod
    def _prune_ein_cache(cache_dir, key_hash):
        """Delete cached Ein files of other keys (Beam.py:5566-5581)."""
        if not os.path.isdir(cache_dir):
            return
        for name in os.listdir(cache_dir):
            if name.startswith("ein_chunk_") and name.endswith(".npz") and key_hash not in name:
                try:
                    os.remove(os.path.join(cache_dir, name))
                except OSError:
                    pass

    def _depth_ein(self, dein, sample, chunk_ids, sizes, d_pos, d_cod, fwd, rot, trans, E0,
                   cache_dir, recompute):
        """Per-atom depth-dependent Ein for one pose, through the Ein cache
        when ``cache_dir`` is set (one ``ein_chunk_<cid>_<hash>.npz`` per
        chunk, as the reference writes them, Beam.py:5021-5044)."""
        s_min, s_max = dein.bounds(d_pos, rot, trans)
        key = paths = None
        if cache_dir:
            key = self._ein_cache_key(sample, rot, trans, s_min, s_max, self._ein_kernel_radius)
            self._prune_ein_cache(cache_dir, key)
            os.makedirs(cache_dir, exist_ok=True)
            paths = [os.path.join(cache_dir, f"ein_chunk_{cid}_{key}.npz") for cid in chunk_ids]
            if not recompute and all(os.path.isfile(p) for p in paths):
                parts = []
                for p in paths:
                    with np.load(p) as z:
                        parts.append(np.asarray(z["ein"], dtype=np.complex64))
                host = np.concatenate(parts) if parts else np.zeros(0, np.complex64)
                if host.shape[0] == d_pos.shape[0]:
                    self._last_ein_source = "cache"
                    return torch.from_numpy(host).to(d_pos.device)
        dein.column_map(d_pos, d_cod, fwd, rot, trans, self._beam_du, self._beam_dv,
                        self._beam_uc, self._beam_vc, self._wavelength * 1e10,
                        self._ein_kernel_radius)
        amp = dein.sample(d_pos, rot, trans, E0, self._beam_du, self._beam_dv, self._beam_uc,
                          self._beam_vc, s_min, s_max)
        self._last_ein_source = "computed"
        if paths is not None:
            host = amp.cpu().numpy()
            off = 0
            for p, n in zip(paths, sizes):
                np.savez_compressed(p, ein=host[off:off + n])
                off += n
        return amp

    @property
    def _last_stats(self):
        """(largest tile radius [wavelengths], pixel geometry) of the last
        staged sum: read from its workspace on first use (waits for it)."""
        s = self._last_staged
        return (s.radius_wl, tuple(s.pixel_geom))

    def _kin_field(self, sample, coords_A, shape, stage, frames, remove_forward,
                   apply_polarization, spherical_decay, out_dtype=torch.complex128,
                   analyser=None, use_depth_ein=False, ein_cache_dir=None,
                   recompute_cache=False):
        """Sum every chunk of ``sample`` for the poses ``frames`` = (rot, trans).

        All chunks go into one device array (pinned host buffer, copies on a
        side stream overlapping the loading of the next chunk), are staged
        once and summed in one launch for all poses.  The incident amplitude
        is sampled at every pose (Beam.py:6279-6290); where it differs
        between poses (a beam profile the sample moves through), each pose is
        re-staged with its own amplitudes (no re-sort) and summed alone.
        ``analyser`` = (kind, axis_m, accept_rad, darwin_hw_rad).  Returns a
        (F, Nz, Ny) device tensor.
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
        ana = (0, (1.0, 0.0, 0.0), 0.0, 0.0) if analyser is None else analyser
        E0 = torch.from_numpy(self._beam_E0_map).to(dev)
        compute = torch.cuda.current_stream(dev)
        copy = self.__dict__.setdefault("_copy_stream", torch.cuda.Stream(dev))
        # chunk groups of at most MAX_GROUP_ATOMS atoms (one group normally)
        groups, cur, cur_n = [], [], 0
        for cid in range(1, n_chunks + 1):
            n = int(sample.chunk_atom_count(cid))
            if cur and cur_n + n > self.MAX_GROUP_ATOMS:
                groups.append(cur)
                cur, cur_n = [], 0
            cur.append(cid)
            cur_n += n
        groups.append(cur)
        labels, index = [], {}
        staged_prev = None
        if use_depth_ein and len(groups) > 1:
            raise NotImplementedError("depth-dependent Ein over more than "
                                      f"{self.MAX_GROUP_ATOMS} atoms")
        cache_dir = None
        if use_depth_ein:
            cache_dir = ein_cache_dir or (os.path.join(self.directory, "ein_cache")
                                          if self.directory else None)
        for g in groups:
            d_pos, d_cod, counts, ev = self._gather_chunks(sample, g, labels, index, dev, copy)
            compute.wait_event(ev)
            if d_pos.shape[0] == 0:
                continue
            consts = tables.species_constants(labels, self._energy)
            spc_tab = K.species_table(consts)
            counts = np.concatenate([counts, np.zeros(len(labels) - counts.shape[0],
                                                      dtype=np.int64)])[:len(labels)]
            opts = K.kin_opts(k_val, remove_forward, apply_polarization,
                              self._polarization_state(), spherical_decay, path=self._kin_path,
                              analyser=ana[0], analyser_axis=ana[1], accept_rad=ana[2],
                              darwin_hw_rad=ana[3])
            if use_depth_ein:
                dein = K.DepthEin(self._beam_Ny, self._beam_Nz, dev)
                fwd = np.stack([(consts["f0z"] + consts["fp"]).astype(np.float32),
                                consts["fpp"]], axis=1)
                sizes = [int(sample.chunk_atom_count(c)) for c in g]

            def ein(f):
                if use_depth_ein:
                    return self._depth_ein(dein, sample, g, sizes, d_pos, d_cod, fwd, rot[f],
                                           trans[f], E0, cache_dir, recompute_cache)
                return K.ein_bilinear(d_pos, rot[f], trans[f], E0, self._beam_du,
                                      self._beam_dv, self._beam_uc, self._beam_vc)

            amp0 = ein(0)
            same = all(K.same_amplitudes(amp0, ein(f)) for f in range(1, F))
            staged = K.stage(d_pos, d_cod, amp0, counts, k_val, F, Ny, Nz,
                             workspace=self._workspace)
            if (staged_prev is not None and staged.workspace is staged_prev.workspace
                    and staged_prev.pixels_key is not None
                    and staged_prev.counts.shape == staged.counts.shape):
                # pixel tables sit at the head of the workspace and survive staging
                staged.pixels_key = staged_prev.pixels_key
                staged.pixel_geom = staged_prev._pixel_geom
            self._workspace = staged.workspace
            if same:
                K.kin_sum(staged, spc_tab, opts, rot, trans, coords, out)
            else:
                for f in range(F):
                    if f > 0:
                        K.restage(staged, d_pos, ein(f))
                    K.kin_sum(staged, spc_tab, opts, rot[f:f + 1], trans[f:f + 1], coords,
                              out[f:f + 1])
            self._last_path = staged.path_used
            self._last_staged = staged
            staged_prev = staged
            del d_pos, d_cod, amp0
        return out

    # ------------------------------------------------------- public calls
    def interact_beam_gpu(self, sample, measurement_positions, measurement_shape, stage,
                          remove_forward=False, use_depth_ein=False, ein_cache_dir=None,
                          recompute_cache=False, apply_polarization=False, spherical_decay=True,
                          analyser_mode="off", analyser_acceptance_angle_rad=0.0,
                          analyser_darwin_halfwidth_rad=0.0):
        Ny, Nz = measurement_shape
        rot = np.asarray(stage.rotation, dtype=np.float32).reshape(1, 3, 3)
        trans = np.asarray(stage.translation, dtype=np.float32).reshape(1, 3)
        mp = measurement_positions
        if isinstance(mp, torch.Tensor):
            mp = mp.detach().cpu().numpy()
        mp = np.asarray(mp, dtype=np.float64)
        out = self._kin_field(sample, mp, (Ny, Nz), stage, (rot, trans), remove_forward,
                              apply_polarization, spherical_decay,
                              analyser=self._analyser(mp, Ny, Nz, analyser_mode,
                                                      analyser_acceptance_angle_rad,
                                                      analyser_darwin_halfwidth_rad),
                              use_depth_ein=use_depth_ein, ein_cache_dir=ein_cache_dir,
                              recompute_cache=recompute_cache)
        return out[0].to(torch.complex64).cpu().numpy()

    @staticmethod
    def _analyser(mp, Ny, Nz, mode, accept_rad, darwin_hw_rad):
        """(kind, axis [m], accept, hw) of the reference's analyser: its axis
        is the direction of the detector's centre pixel (Ny // 2, Nz // 2)
        (Beam.py:6128-6143)."""
        kind = K.analyser_kind(mode)
        if kind == 0:
            return None
        c = (Nz // 2) * Ny + Ny // 2
        axis = tuple(float(v) / 1e10 for v in np.asarray(mp)[:, c])
        return (kind, axis, float(accept_rad), float(darwin_hw_rad))

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
        at every pose; where it is the same for all of them, the poses run in
        one launch.
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

