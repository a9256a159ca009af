"""Chunked samples (in memory or in the reference's chunk files) and the
synthetic crystals of the workloads.

The kinematic path consumes a sample through these members, with the
reference's meaning (Sample.py:351-420, 543-676, 735-868):

* ``chunk_total`` -- number of chunks, indexed 1..chunk_total;
* ``chunk_atom_count(cid)``;
* ``load_chunk_positions(cid, use_gpu=False, raw=False)`` -- (N, 3) Angstrom
  (a copy, with the configured thermal expansion and displacements applied on
  load unless ``raw``);
* ``load_chunk_species(cid, use_gpu=False, codes=False)`` -- labels, or
  ``(codes, label table)`` with ``codes=True``.

On disk a sample is the reference's directory layout: ``atomic_positions_<cid>.npy``
(N, 3) positions, ``atomic_species_<cid>.npy`` (uint8/uint16 codes into the
``species_labels`` table of ``sample_metadata.json``, or the labels
themselves), and ``sample_metadata.json`` (``chunk_total``, ``species_labels``,
``dimensions``, ``offset``, ``rotation``, ...).  ``sample.from_directory``
reads one; ``write_chunk_positions`` / ``write_chunk_species`` /
``write_sample_metadata`` write one.  Files are read with ``numpy.load``
(no pickles).

Thermal displacements on load follow ``apply_temperature`` (Sample.py:
2084-2338): a generator keyed on ``SeedSequence([seed, chunk])``; Gaussian
(sigma in Angstrom, drawn in the positions' dtype), Einstein (sigma is the
temperature: <x^2> = hbar / (2 m omega) coth(hbar omega / 2 k_B T), omega =
k_B theta_E / hbar, per species where configured) and Debye (<u^2> =
9 hbar^2 / (m k_B theta_D) [phi(theta_D / T) + 1/4]) modes, each clipped to
+-max_displacement; thermal expansion (Sample.py:1977-2070) scales positions
about the sample centre first.

Generation (lattices, orientation, defect displacement fields) replaces the
reference's CIF/pymatgen machinery, which is out of scope; it produces the
synthetic crystals that the workloads name.
"""
import json
import os

import numpy as np

FCC_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
PEROVSKITE_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [0.5, 0.5, 0],
                             [0.5, 0, 0.5], [0, 0.5, 0.5]])

K_B = 1.380649e-23
HBAR = 1.054571817e-34
AMU_KG = 1.66053906660e-27
POSITIONS_FILE = "atomic_positions.npy"
SPECIES_FILE = "atomic_species.npy"
METADATA_FILE = "sample_metadata.json"


def _chunk_name(default, cid):
    base, ext = os.path.splitext(default)
    return f"{base}_{int(cid)}{ext}"


class sample:
    """A chunked sample: in memory (lists of chunk arrays) or on disk
    (``sample.from_directory``)."""

    def __init__(self, chunks_pos=None, chunks_codes=None, labels=(), directory=""):
        chunks_pos = [] if chunks_pos is None else chunks_pos
        chunks_codes = [] if chunks_codes is None else chunks_codes
        if len(chunks_pos) != len(chunks_codes):
            raise ValueError("positions and species chunk lists differ in length")
        self._pos = [np.ascontiguousarray(p, dtype=np.float32) for p in chunks_pos]
        self._codes = [np.ascontiguousarray(c, dtype=np.uint8) for c in chunks_codes]
        for p, c in zip(self._pos, self._codes):
            if p.ndim != 2 or p.shape[1] != 3 or p.shape[0] != c.shape[0]:
                raise ValueError("chunk positions must be (N, 3) with N species codes")
        self._species_labels = [str(x) for x in labels]
        self.directory = directory
        self._on_disk = False
        self._chunk_total = len(self._pos)
        self._counts = None
        self._dimensions = None
        self._offset = None
        self._rotation = None
        self._sample_type = None
        self.enable_temp = False
        self.temp_params = None

    # -------------------------------------------------------- on disk
    @classmethod
    def from_directory(cls, directory):
        """A sample over the chunk files in ``directory`` (read lazily)."""
        s = cls(directory=str(directory))
        s._on_disk = True
        s.read_sample_metadata()
        return s

    def read_sample_metadata(self):
        """Restore chunk_total, the label table and the geometry fields from
        sample_metadata.json (Sample.py:351-386)."""
        path = os.path.join(self.directory, METADATA_FILE)
        if not os.path.isfile(path):
            raise FileNotFoundError(f"No JSON metadata file found at {path}")
        with open(path, "r") as f:
            meta = json.load(f)
        if meta.get("dimensions") is not None:
            self._dimensions = np.array(meta["dimensions"], dtype=np.float32)
        if meta.get("offset") is not None:
            self._offset = np.array(meta["offset"], dtype=np.float32)
        if meta.get("rotation") is not None:
            self._rotation = np.array(meta["rotation"], dtype=np.float32)
        if meta.get("chunk_total") is not None:
            self._chunk_total = int(meta["chunk_total"])
        self._sample_type = meta.get("sample_type")
        self._species_labels = [str(x) for x in meta.get("species_labels", [])]
        self._counts = None

    def write_sample_metadata(self, override_directory=None):
        """sample_metadata.json with the fields the reference writes
        (Sample.py:636-676)."""
        meta = {
            "dimensions": None if self._dimensions is None else np.asarray(self._dimensions).tolist(),
            "offset": None if self._offset is None else np.asarray(self._offset).tolist(),
            "rotation": None if self._rotation is None else np.asarray(self._rotation).tolist(),
            "chunk_total": int(self.chunk_total),
            "sample_type": self._sample_type,
            "alloy_species": None,
            "alloy_concentrations": None,
            "sample_shape": "box",
            "modifications": [],
            "species_labels": list(self._species_labels),
        }
        d = override_directory if override_directory is not None else self.directory
        with open(os.path.join(d, METADATA_FILE), "w") as f:
            json.dump(meta, f, indent=4)

    def write_chunk_positions(self, data, chunk_num, override_directory=None):
        d = override_directory if override_directory is not None else self.directory
        np.save(os.path.join(d, _chunk_name(POSITIONS_FILE, chunk_num)), np.asarray(data))

    def write_chunk_species(self, data, chunk_num, override_directory=None):
        """Codes (uint8, or uint16 beyond 256 labels) into the label table,
        extending it with new labels, as the reference writes them in the
        sample's own directory (Sample.py:570-634)."""
        data = np.asarray(data)
        if data.dtype.kind in "iu":
            if data.size and int(data.max()) >= len(self._species_labels):
                raise ValueError("species codes exceed the sample's label table")
            codes = data
        else:
            labels, inv = np.unique(data.astype(str), return_inverse=True)
            lut = np.empty(labels.size, dtype=np.int64)
            for i, lab in enumerate(labels):
                if str(lab) not in self._species_labels:
                    self._species_labels.append(str(lab))
                lut[i] = self._species_labels.index(str(lab))
            codes = lut[inv.reshape(-1)]
        dt = np.uint8 if len(self._species_labels) <= 256 else np.uint16
        d = override_directory if override_directory is not None else self.directory
        np.save(os.path.join(d, _chunk_name(SPECIES_FILE, chunk_num)), codes.astype(dt))

    def write_chunks(self, directory):
        """Write every chunk (raw positions, species codes) and the metadata
        into ``directory`` in the reference's layout."""
        os.makedirs(directory, exist_ok=True)
        for cid in range(1, self.chunk_total + 1):
            self.write_chunk_positions(self.load_chunk_positions(cid, raw=True), cid, directory)
            codes, _ = self.load_chunk_species(cid, codes=True)
            self.write_chunk_species(codes, cid, directory)
        self.write_sample_metadata(directory)

    def chunk_position_path(self, cid):
        """Path of a chunk's positions file (on-disk samples), else None."""
        if not self._on_disk:
            return None
        return os.path.join(self.directory, _chunk_name(POSITIONS_FILE, cid))

    # ----------------------------------------------------------- interface
    @property
    def chunk_total(self):
        return self._chunk_total

    def chunk_atom_count(self, chunk_number):
        if not self._on_disk:
            return int(self._pos[chunk_number - 1].shape[0])
        if self._counts is None:
            self._counts = {}
        n = self._counts.get(chunk_number)
        if n is None:
            arr = np.load(self.chunk_position_path(chunk_number), mmap_mode="r")
            n = int(arr.shape[0])
            self._counts[chunk_number] = n
        return n

    @property
    def atom_count(self):
        return int(sum(self.chunk_atom_count(c) for c in range(1, self.chunk_total + 1)))

    @property
    def offset(self):
        return np.zeros(3) if self._offset is None else self._offset

    @property
    def dimensions(self):
        return self._dimensions

    def set_temperature(self, distribution="gaussian", sigma=0.25, max_displacement=1,
                        seed=40):
        dist = str(distribution).lower()
        if dist not in ("gaussian", "normal", "einstein", "temperature", "kelvin", "debye"):
            raise ValueError(f"Unknown distribution: {distribution}")
        self.enable_temp = True
        self.temp_params = [str(distribution), float(sigma), max_displacement, seed]

    def set_temperature_einstein(self, T_K, mass_amu=None, theta_E_K=None,
                                 species_mass_amu=None, species_theta_E_K=None,
                                 max_displacement=None, seed=40):
        """Einstein-model displacements (Sample.py:1868-1929)."""
        self.enable_temp = True
        self.temp_params = ["einstein", float(T_K), max_displacement, seed]
        if mass_amu is not None:
            self._temp_mass_amu = float(mass_amu)
        if theta_E_K is not None:
            self._temp_theta_E_K = float(theta_E_K)
        if species_mass_amu is not None:
            self._temp_species_mass_amu = dict(species_mass_amu)
        if species_theta_E_K is not None:
            self._temp_species_theta_E_K = dict(species_theta_E_K)
        if not hasattr(self, "_position_unit_in_m"):
            self._position_unit_in_m = 1.0e-10

    def set_temperature_debye(self, T_K, mass_amu=None, theta_D_K=None, max_displacement=None,
                              seed=40):
        """Debye-model displacements (Sample.py:1931-1975)."""
        self.enable_temp = True
        self.temp_params = ["debye", float(T_K), max_displacement, seed]
        if mass_amu is not None:
            self._temp_mass_amu = float(mass_amu)
        if theta_D_K is not None:
            self._temp_theta_D_K = float(theta_D_K)
        if not hasattr(self, "_position_unit_in_m"):
            self._position_unit_in_m = 1.0e-10

    def set_thermal_expansion(self, alpha=None, alpha_xyz=None, T_ref=300.0):
        """Scale positions about the sample centre by 1 + alpha (T - T_ref) on
        load (Sample.py:1977-2013)."""
        self._thermal_expansion_enabled = True
        self._thermal_expansion_T_ref = float(T_ref)
        if alpha_xyz is not None:
            self._thermal_expansion_alpha_xyz = np.array(alpha_xyz, dtype=np.float64)
            self._thermal_expansion_alpha = None
        elif alpha is not None:
            self._thermal_expansion_alpha = float(alpha)
            self._thermal_expansion_alpha_xyz = None

    def apply_thermal_expansion(self, positions, T_K):
        if not getattr(self, "_thermal_expansion_enabled", False):
            return positions
        dT = T_K - getattr(self, "_thermal_expansion_T_ref", 300.0)
        center = np.asarray(self.offset, dtype=np.float64)
        a_xyz = getattr(self, "_thermal_expansion_alpha_xyz", None)
        a = getattr(self, "_thermal_expansion_alpha", None)
        if a_xyz is not None:
            return center + (positions - center) * (1.0 + np.asarray(a_xyz) * dT)
        if a is not None:
            return center + (positions - center) * (1.0 + a * dT)
        return positions

    @staticmethod
    def _thermal_rng(seed, chunk_number):
        root = np.random.SeedSequence().entropy if seed is None else int(seed)
        key = [root] if chunk_number is None else [root, int(chunk_number)]
        return np.random.default_rng(np.random.SeedSequence(key))

    def _atom_masses_theta(self, chunk_number, n):
        m_def = getattr(self, "_temp_mass_amu", 28.0)
        th_def = getattr(self, "_temp_theta_E_K", 300.0)
        per_species = (getattr(self, "_temp_species_mass_amu", None) is not None
                       and getattr(self, "_temp_species_theta_E_K", None) is not None
                       and chunk_number is not None)
        if not per_species:
            return np.full(n, m_def), np.full(n, th_def)
        species = np.asarray(self.load_chunk_species(chunk_number))
        if species.shape[0] != n:
            raise ValueError(f"chunk {chunk_number}: {species.shape[0]} species for {n} positions")
        labels, inv = np.unique(species.astype(str), return_inverse=True)
        mt = np.array([self._temp_species_mass_amu.get(str(k), m_def) for k in labels])
        tt = np.array([self._temp_species_theta_E_K.get(str(k), th_def) for k in labels])
        return mt[inv.reshape(-1)], tt[inv.reshape(-1)]

    @staticmethod
    def debye_phi(x):
        """phi(x) = (3 / x^4) int_0^x t^3 / (e^t - 1) dt, with the large-x limit
        pi^4 / 15 for the integral beyond x = 50 (Sample.py:2276-2300)."""
        from scipy.integrate import quad

        def f(t):
            return t * t if t < 1e-10 else t * t * t / (np.exp(t) - 1.0)
        val = (np.pi ** 4) / 15.0 if x > 50.0 else quad(f, 0, x, limit=100)[0]
        return (3.0 / x ** 4) * val

    def thermal_sigma(self, distribution, T_or_sigma, chunk_number=None, n=1):
        """Per-atom displacement sigma [position units] of a thermal mode."""
        d = str(distribution).lower()
        if d in ("gaussian", "normal"):
            return np.full(n, float(T_or_sigma))
        unit = getattr(self, "_position_unit_in_m", 1e-10)
        T = float(T_or_sigma)
        if d in ("einstein", "temperature", "kelvin"):
            m_amu, th = self._atom_masses_theta(chunk_number, n)
            m = np.asarray(m_amu, dtype=np.float64) * AMU_KG
            omega = K_B * np.asarray(th, dtype=np.float64) / HBAR
            if T <= 0.0:
                coth = np.ones_like(omega)
            else:
                z = HBAR * omega / (2.0 * K_B * T)
                coth = np.where(z < 1e-6, 1.0 / z + z / 3.0, 1.0 / np.tanh(z))
            return np.sqrt(HBAR / (2.0 * m * omega) * coth) / unit
        if d == "debye":
            m = getattr(self, "_temp_mass_amu", 28.0) * AMU_KG
            thD = getattr(self, "_temp_theta_D_K", 300.0)
            phi = 0.0 if T <= 0.0 else self.debye_phi(thD / T)
            msd = 9.0 * HBAR * HBAR / (m * K_B * thD) * (phi + 0.25)
            return np.full(n, np.sqrt(msd) / unit)
        raise ValueError(f"Unknown distribution: {distribution}")

    def apply_temperature(self, positions, distribution="gaussian", sigma=0.25,
                          max_displacement=1, seed=40, chunk_number=None):
        """Random displacements of one chunk's positions (Sample.py:2084-2318)."""
        rng = self._thermal_rng(seed, chunk_number)
        d = str(distribution).lower()
        if d in ("gaussian", "normal"):
            disp = rng.standard_normal(size=positions.shape, dtype=positions.dtype)
            disp *= float(sigma)
        elif d in ("einstein", "temperature", "kelvin", "debye"):
            sig = self.thermal_sigma(d, sigma, chunk_number, int(positions.shape[0]))
            rand = rng.standard_normal(size=positions.shape)
            disp = rand * (sig.reshape(-1, 1) if d != "debye" else sig[0])
            disp = disp.astype(positions.dtype, copy=False)
        else:
            raise ValueError(f"Unknown distribution: {distribution}")
        if (max_displacement is not None) and (max_displacement > 0.0):
            np.clip(disp, -max_displacement, max_displacement, out=disp)
        return positions + disp

    def _raw_positions(self, chunk_number):
        if self._on_disk:
            return np.load(self.chunk_position_path(chunk_number))
        return np.array(self._pos[chunk_number - 1], copy=True)

    def pin_memory(self):
        """Move the in-memory chunks into page-locked host memory (one block
        for the positions, one for the species codes) so that the drop-in
        copies them to the device by DMA, without a host copy.  Chunk files
        are not affected.  Returns the sample."""
        if self._on_disk or not self._pos or getattr(self, "_pinned", None) is not None:
            return self
        import torch
        n = sum(int(p.shape[0]) for p in self._pos)
        pos = torch.empty((max(n, 1), 3), dtype=torch.float32, pin_memory=True)
        cod = torch.empty(max(n, 1), dtype=torch.uint8, pin_memory=True)
        pn, cn = pos.numpy(), cod.numpy()
        views_p, views_c, counts = [], [], []
        off = 0
        for p, c in zip(self._pos, self._codes):
            k = int(p.shape[0])
            pn[off:off + k] = p
            cn[off:off + k] = c
            views_p.append(pn[off:off + k])
            views_c.append(cn[off:off + k])
            counts.append(np.bincount(c, minlength=max(1, len(self._species_labels))))
            off += k
        self._pos, self._codes = views_p, views_c
        self._pinned = (pos, cod, counts)
        return self

    def pinned_chunk(self, chunk_number):
        """(positions (N, 3) float32, codes (N,) uint8 -- pinned torch
        tensors, per-species counts of the chunk's codes) when the sample is
        pinned and nothing applies on load; None otherwise."""
        pin = getattr(self, "_pinned", None)
        if pin is None or self.enable_temp:
            return None
        p, c = self._pos[chunk_number - 1], self._codes[chunk_number - 1]
        import torch
        return torch.from_numpy(p), torch.from_numpy(c), pin[2][chunk_number - 1]

    def chunk_positions_view(self, chunk_number):
        """The chunk's positions without a copy (the stored array, or a
        memory map of the chunk file) when nothing applies on load (no
        thermal displacements or expansion); None otherwise."""
        if self.enable_temp:
            return None
        if self._on_disk:
            return np.load(self.chunk_position_path(chunk_number), mmap_mode="r")
        return self._pos[chunk_number - 1]

    def load_chunk_positions(self, chunk_number, use_gpu=False, raw=False):
        pos = self._raw_positions(chunk_number)
        if self.enable_temp and not raw:
            dist, sigma, maxd, seed = self.temp_params
            T_K = (getattr(self, "_thermal_expansion_T_ref", 300.0)
                   if str(dist).lower() in ("gaussian", "normal") else float(sigma))
            if getattr(self, "_thermal_expansion_enabled", False):
                pos = self.apply_thermal_expansion(pos, T_K)
            pos = self.apply_temperature(pos, dist, sigma, maxd, seed, chunk_number)
        return pos

    def load_chunk_species(self, chunk_number, use_gpu=False, codes=False):
        if self._on_disk:
            arr = np.load(os.path.join(self.directory, _chunk_name(SPECIES_FILE, chunk_number)))
        else:
            arr = self._codes[chunk_number - 1]
        labels = np.asarray(self._species_labels, dtype=str)
        if arr.dtype.kind in "iu":
            if arr.size and int(arr.max()) >= labels.size:
                raise ValueError(f"species codes in chunk {chunk_number} exceed the label "
                                 f"table ({labels.size} labels); read_sample_metadata() may "
                                 "be missing")
            return (arr, labels) if codes else labels[arr]
        # labels stored directly
        if codes:
            lab, inv = np.unique(arr.astype(str), return_inverse=True)
            return inv.reshape(-1).astype(np.uint8 if lab.size <= 256 else np.uint16), lab
        return np.array(arr, copy=True)

    @property
    def species_labels(self):
        return np.asarray(self._species_labels, dtype=str)


# --------------------------------------------------------------- lattices
def lattice(a, cells, basis):
    """Positions (float64 Angstrom) of a cubic lattice, centred on the origin.

    Atom order: cells in C order over (ix, iy, iz), basis inner.
    Returns (positions, basis_index).
    """
    nx, ny, nz = (int(c) for c in cells)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cell = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    b = np.asarray(basis, dtype=np.float64)
    pos = (cell[:, None, :] + b[None, :, :]).reshape(-1, 3) * float(a)
    centre = 0.5 * np.array([nx, ny, nz], dtype=np.float64) * float(a)
    pos -= centre
    bidx = np.tile(np.arange(b.shape[0]), cell.shape[0])
    return pos, bidx


def rotation_to(v_from, v_to):
    """Smallest rotation taking direction v_from to v_to (float64 3x3)."""
    a = np.asarray(v_from, dtype=np.float64)
    b = np.asarray(v_to, dtype=np.float64)
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    v = np.cross(a, b)
    c = float(np.dot(a, b))
    s = float(np.linalg.norm(v))
    if s < 1e-15:
        if c > 0:
            return np.eye(3)
        # 180 degrees about any axis normal to a
        t = np.array([1.0, 0, 0]) if abs(a[0]) < 0.9 else np.array([0, 1.0, 0])
        ax = np.cross(a, t)
        ax /= np.linalg.norm(ax)
        return 2 * np.outer(ax, ax) - np.eye(3)
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + K + K @ K * ((1 - c) / (s * s))


def bragg_angle(d_spacing, wavelength_m):
    """theta_B [rad] for a d-spacing in Angstrom."""
    return float(np.arcsin(wavelength_m * 1e10 / (2.0 * d_spacing)))


def orient_for_reflection(hkl, theta_b):
    """Rotation putting the reciprocal vector (hkl) along q_hat.

    For a +x beam and a detector at 2theta in the x-z plane (Detector.py:
    700-705) the scattering vector is along (-sin theta, 0, cos theta).
    """
    target = np.array([-np.sin(theta_b), 0.0, np.cos(theta_b)])
    return rotation_to(np.asarray(hkl, dtype=np.float64), target)


def split_chunks(pos, codes, n_chunks):
    """Split contiguous atom ranges into ``n_chunks`` chunks."""
    n = pos.shape[0]
    edges = np.linspace(0, n, int(n_chunks) + 1).round().astype(np.int64)
    return ([pos[edges[i]:edges[i + 1]] for i in range(n_chunks)],
            [codes[edges[i]:edges[i + 1]] for i in range(n_chunks)])


def block_chunks(pos, codes, a, cells, blocks):
    """Chunks as spatial blocks of cells (blocks = (bx, by, bz) per axis)."""
    nx, ny, nz = cells
    cell_idx = np.floor((pos + 0.5 * np.array(cells) * a) / a + 1e-6).astype(np.int64)
    cell_idx = np.clip(cell_idx, 0, np.array(cells) - 1)
    bsz = np.array([-(-nx // blocks[0]), -(-ny // blocks[1]), -(-nz // blocks[2])])
    bi = cell_idx // bsz
    key = (bi[:, 0] * blocks[1] + bi[:, 1]) * blocks[2] + bi[:, 2]
    order = np.argsort(key, kind="stable")
    key_s = key[order]
    nb = blocks[0] * blocks[1] * blocks[2]
    bounds = np.searchsorted(key_s, np.arange(nb + 1))
    return ([pos[order[bounds[i]:bounds[i + 1]]] for i in range(nb)],
            [codes[order[bounds[i]:bounds[i + 1]]] for i in range(nb)])


# --------------------------------------------------------- displacements
def screw_dislocation_uz(pos, b=4.0, centre=(0.0, 0.0)):
    """u_z = b * atan2(y - y0, x - x0) / (2 pi) (screw line along z)."""
    return b * np.arctan2(pos[:, 1] - centre[1], pos[:, 0] - centre[0]) / (2 * np.pi)


def stacking_fault_shift(pos_crystal, a, plane_offset=0.0):
    """Intrinsic fault on a (111) plane: atoms above shift by a/6 [1 1 -2].

    ``pos_crystal`` are crystal-frame positions; the plane passes through
    n . r = plane_offset with n = [111]/sqrt(3).  Mirrors the one-fault case
    of the reference's rule (Defects.py:3951-3990): atoms with the plane
    strictly below them get the Burgers vector.
    """
    n = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    above = (pos_crystal @ n) > plane_offset
    shift = np.array([1.0, 1.0, -2.0]) * (a / 6.0)
    out = pos_crystal.copy()
    out[above] += shift
    return out


def gaussian_bump(pos, r0, width, amplitude, direction):
    """u(r) = amplitude * exp(-|r - r0|^2 / (2 width^2)) * e_hat."""
    e = np.asarray(direction, dtype=np.float64)
    e = e / np.linalg.norm(e)
    d2 = ((pos - np.asarray(r0)[None, :]) ** 2).sum(axis=1)
    return amplitude * np.exp(-d2 / (2.0 * width * width))[:, None] * e[None, :]
