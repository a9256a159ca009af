"""Differentiable kinematic field for displacement-field fitting.

``KinematicModel`` holds a sample (undisplaced sample-frame positions,
species), a beam, a detector and a list of sample poses.  ``field(u)`` returns
the complex field A (F, Nz, Ny) complex128 for per-atom displacements u
(N, 3) float32 [Angstrom, sample frame] as a torch autograd node: the forward
is mosaic_kin_sum on the displaced positions, the backward the hand-written
adjoint kernel (mosaic_kin_adjoint), which recomputes phases instead of
storing them.  Form factors, polarization and the incident amplitudes are
held fixed with respect to u (their dependence on the position is of order
|u| / distance).  The atoms are sorted into tiles once; each forward
re-stages the displaced atoms into the same tiles (mosaic_kin_restage).
"""
import numpy as np
import torch

from . import kinematic as K
from . import tables


class _KinField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, model):
        model._forward(u)
        ctx.model = model
        ctx.version = model._version
        ctx.pos = model._pos
        ctx.n = u.shape[0]
        return model._out.clone()

    @staticmethod
    def backward(ctx, grad_A):
        m = ctx.model
        if m._version != ctx.version:
            # a later forward re-staged the workspace: restore this one's tiles
            K.restage(m._staged, ctx.pos, m.amp)
            m._version = ctx.version
            m._pos = ctx.pos
        grad = torch.zeros((ctx.n, 3), dtype=torch.float32, device=grad_A.device)
        K.kin_adjoint(m._staged, m.k, m.rot, m.trans, grad_A, grad, path=m.adjoint_path)
        return grad, None


class KinematicModel:
    def __init__(self, positions, codes, labels, beam, detector, rot, trans,
                 apply_polarization=True, remove_forward=False, device=None,
                 adjoint_path="auto"):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.adjoint_path = adjoint_path
        self.pos0 = torch.as_tensor(np.ascontiguousarray(positions, dtype=np.float32),
                                    device=self.device)
        self.codes = torch.as_tensor(np.ascontiguousarray(codes, dtype=np.uint8),
                                     device=self.device)
        self.counts = np.bincount(np.asarray(codes, dtype=np.int64),
                                  minlength=len(labels)).astype(np.int64)
        self.rot = np.asarray(rot, dtype=np.float32).reshape(-1, 3, 3)
        self.trans = np.asarray(trans, dtype=np.float32).reshape(-1, 3)
        self.F = self.rot.shape[0]
        self.Ny, self.Nz = detector.shape
        self.k = 2.0 * np.pi / beam.wavelength
        self.spc = K.species_table(tables.species_constants(list(labels), beam.energy))
        self.opts = K.kin_opts(self.k, remove_forward, apply_polarization,
                               beam._polarization_state(), False, path="auto")
        self.coords = torch.as_tensor(np.ascontiguousarray(detector.pixel_coordinates / 1e10),
                                      device=self.device)
        E0 = torch.as_tensor(beam._beam_E0_map, device=self.device)
        amps = [K.ein_bilinear(self.pos0, self.rot[f], self.trans[f], E0, beam._beam_du,
                               beam._beam_dv, beam._beam_uc, beam._beam_vc)
                for f in range(self.F)]
        if not all(K.same_amplitudes(amps[0], a) for a in amps[1:]):
            raise NotImplementedError(
                "KinematicModel holds one incident amplitude per atom: the beam profile must "
                "give the same amplitudes at every pose (e.g. a uniform beam covering the "
                "sample)")
        self.amp = amps[0]
        self._out = torch.zeros((self.F, self.Nz, self.Ny), dtype=torch.complex128,
                                device=self.device)
        # one staged sample for the model's lifetime: the first forward sorts
        # the atoms into tiles, later forwards re-stage the moved atoms into the
        # same tiles (no sort, no allocation); the adjoint's scratch stays with it
        self._staged = None
        self._pos = None
        self._version = 0
        self.pixel_geom = None

    def _forward(self, u):
        pos = (self.pos0 + u.detach().to(torch.float32) if u is not None
               else self.pos0).contiguous()
        if self._staged is None:
            self._staged = K.stage(pos, self.codes, self.amp, self.counts, self.k, self.F,
                                   self.Ny, self.Nz)
        else:
            K.restage(self._staged, pos, self.amp)
        self._version += 1
        self._pos = pos
        staged = self._staged
        self._out.zero_()
        K.kin_sum(staged, self.spc, self.opts, self.rot, self.trans, self.coords, self._out)
        self.path_used = staged.path_used
        if self.pixel_geom is None:             # the detector's, fixed: read once
            self.pixel_geom = tuple(staged.pixel_geom)
        return staged

    def field(self, u):
        """Complex field (F, Nz, Ny) complex128, differentiable in u."""
        return _KinField.apply(u, self)
