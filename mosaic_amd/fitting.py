"""Differentiable kinematic field for displacement-field fitting.

``KinematicModel`` holds a sample (undisplaced sample-frame positions,
species), a beam, a detector and a list of sample poses.  ``field(u)`` returns
the complex field A (F, Nz, Ny) complex128 for per-atom displacements u
(N, 3) float32 [Angstrom, sample frame] as a torch autograd node: the forward
is mosaic_kin_sum on the displaced positions, the backward the hand-written
adjoint kernel (mosaic_kin_adjoint), which recomputes phases instead of
storing them.  Form factors, polarization and the incident amplitudes are
held fixed with respect to u (their dependence on the position is of order
|u| / distance).
"""
import numpy as np
import torch

from . import kinematic as K
from . import tables


class _KinField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, model):
        staged = model._forward(u)
        ctx.model = model
        ctx.staged = staged
        ctx.n = u.shape[0]
        return model._out.clone()

    @staticmethod
    def backward(ctx, grad_A):
        m = ctx.model
        grad = torch.zeros((ctx.n, 3), dtype=torch.float32, device=grad_A.device)
        K.kin_adjoint(ctx.staged, m.k, m.rot, m.trans, grad_A, grad, path=m.adjoint_path)
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
        self.amp = K.ein_bilinear(self.pos0, self.rot[0], self.trans[0], E0, beam._beam_du,
                                  beam._beam_dv, beam._beam_uc, beam._beam_vc)
        self._out = torch.zeros((self.F, self.Nz, self.Ny), dtype=torch.complex128,
                                device=self.device)

    def _forward(self, u):
        pos = self.pos0 + u.detach().to(torch.float32) if u is not None else self.pos0
        # a fresh workspace per forward: the backward of this call reads it
        staged = K.stage(pos.contiguous(), self.codes, self.amp, self.counts, self.k, self.F,
                         self.Ny, self.Nz)
        self._out.zero_()
        K.kin_sum(staged, self.spc, self.opts, self.rot, self.trans, self.coords, self._out)
        self.path_used = staged.path_used
        self.pixel_geom = tuple(staged.pixel_geom)
        return staged

    def field(self, u):
        """Complex field (F, Nz, Ny) complex128, differentiable in u."""
        return _KinField.apply(u, self)
