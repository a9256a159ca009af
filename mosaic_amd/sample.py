"""In-memory chunked samples and the synthetic crystals of the workloads.

The kinematic path consumes a sample through four members, which this class
provides with the reference's meaning (Sample.py:735-868):

* ``chunk_total`` -- number of chunks, indexed 1..chunk_total;
* ``chunk_atom_count(cid)``;
* ``load_chunk_positions(cid, use_gpu=False)`` -- (N, 3) float32 Angstrom
  (a copy, with the configured thermal displacements applied on load);
* ``load_chunk_species(cid, use_gpu=False, codes=False)`` -- labels, or
  ``(uint8 codes, label table)`` with ``codes=True``.

Thermal displacements on load follow ``apply_temperature`` in its Gaussian
mode (Sample.py:2149-2156, 2321-2338): a generator keyed on
``SeedSequence([seed, chunk])``, ``standard_normal`` in the positions'
dtype, times sigma, clipped to +-max_displacement.

Generation (lattices, orientation, defect displacement fields) replaces the
reference's CIF/pymatgen and chunk-file machinery, which is out of scope; it
produces the synthetic crystals that the workloads name.
"""
import numpy as np

FCC_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
PEROVSKITE_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0.5], [0.5, 0.5, 0],
                             [0.5, 0, 0.5], [0, 0.5, 0.5]])


class sample:
    """A sample held in host memory as a list of chunks."""

    def __init__(self, chunks_pos, chunks_codes, labels, directory=""):
        if len(chunks_pos) != len(chunks_codes):
            raise ValueError("positions and species chunk lists differ in length")
        self._pos = [np.ascontiguousarray(p, dtype=np.float32) for p in chunks_pos]
        self._codes = [np.ascontiguousarray(c, dtype=np.uint8) for c in chunks_codes]
        for p, c in zip(self._pos, self._codes):
            if p.ndim != 2 or p.shape[1] != 3 or p.shape[0] != c.shape[0]:
                raise ValueError("chunk positions must be (N, 3) with N species codes")
        self._species_labels = np.asarray(labels, dtype=str)
        self.directory = directory
        self.enable_temp = False
        self.temp_params = None

    # ----------------------------------------------------------- interface
    @property
    def chunk_total(self):
        return len(self._pos)

    def chunk_atom_count(self, chunk_number):
        return int(self._pos[chunk_number - 1].shape[0])

    @property
    def atom_count(self):
        return int(sum(p.shape[0] for p in self._pos))

    def set_temperature(self, distribution="gaussian", sigma=0.25, max_displacement=1,
                        seed=40):
        if str(distribution).lower() not in ("gaussian", "normal"):
            raise NotImplementedError("only Gaussian thermal displacements are implemented")
        self.enable_temp = True
        self.temp_params = [str(distribution), float(sigma), max_displacement, seed]

    def apply_temperature(self, positions, sigma, max_displacement, seed, chunk_number):
        root = np.random.SeedSequence().entropy if seed is None else int(seed)
        key = [root] if chunk_number is None else [root, int(chunk_number)]
        rng = np.random.default_rng(np.random.SeedSequence(key))
        disp = rng.standard_normal(size=positions.shape, dtype=positions.dtype)
        disp *= float(sigma)
        if (max_displacement is not None) and (max_displacement > 0.0):
            np.clip(disp, -max_displacement, max_displacement, out=disp)
        return positions + disp

    def load_chunk_positions(self, chunk_number, use_gpu=False, raw=False):
        pos = np.array(self._pos[chunk_number - 1], copy=True)
        if self.enable_temp and not raw:
            _, sigma, maxd, seed = self.temp_params
            pos = self.apply_temperature(pos, sigma, maxd, seed, chunk_number)
        return pos

    def load_chunk_species(self, chunk_number, use_gpu=False, codes=False):
        arr = self._codes[chunk_number - 1]
        if codes:
            return arr, self._species_labels
        return self._species_labels[arr]

    @property
    def species_labels(self):
        return self._species_labels


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
