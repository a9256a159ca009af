"""Learned surrogate of the kinematic rocking scan, with fallback to the plain sum.

Model: an ensemble of ``members`` networks, each an MLP followed by a
convolutional decoder of 2x2 stride-2 transposed convolutions,

    9 inputs -> 256 -> 512 -> 8x8x64 -> 16x16x64 -> 32x32x32 -> 64x64x16 -> 128x128x1

(SiLU between layers, none at the output).  Every layer is one batched GEMM
(``mosaic_sg_gemm``: BF16 MFMA, FP32 accumulation; the ensemble is the GEMM
batch) plus hand-written elementwise kernels; the backward pass and the Adam
update run on the same kernels.  The output is, with ``target="log"`` (the
default), the normalised log-intensity (ln(I / I0 + 1e-4) - mu) / sd trained
on its mean squared error, or, with ``target="rel"``, the intensity patch
y = I / I0 trained on each image's squared relative L2 error
||y' - y||^2 / ||y||^2.

Defect parameters (9): dislocation core x, y [nm, +-50], Burgers vector
polar/azimuth angles, line direction polar/azimuth angles, hydrostatic strain
(0..1e-3), stacking-fault shift fraction (0..1), rocking angle [deg, +-0.05].
The network sees them as 11 features (``features``): the two directions as
unit vectors, so that the azimuth's wrap at 2 pi is no discontinuity.

Training images come from the GPU kinematic sum itself
(``make_training_set``): the defect field is applied on the device
(``mosaic_defect_positions``), the crystal staged and summed at its rocking
angle.  At scan time (``SurrogateScan``) the plain sum computes anchor frames (the
first and last angle), which judge the network and correct each member's
prediction; the other frames come from the corrected ensemble.  Frames on
which the members disagree beyond a limit fall back to the plain sum, and
so do frames that the plain sum judges: an exact probe of each frame at a
few dozen importance-sampled pixels estimates its relative error, and
frames whose estimate's upper bound exceeds the threshold are recomputed.  Inference runs fused (``infer``: two
launches on the 32x32x16 BF16 matrix cores).
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from . import kinematic as K
from . import sg_ops as S
from . import tables
from ._lib import check

N_IN = 9            # defect parameters, rocking angle last
N_FEAT = 11         # network inputs (``features``)
LOG_EPS = 1e-4
A_AU = 4.08
NU_AU = 0.44

# ranges of the 9 inputs (low, high)
RANGES = np.array([
    [-50.0, 50.0],        # core x [nm]
    [-50.0, 50.0],        # core y [nm]
    [0.0, math.pi],       # Burgers polar angle
    [0.0, 2 * math.pi],   # Burgers azimuth
    [0.0, math.pi],       # line polar angle
    [0.0, 2 * math.pi],   # line azimuth
    [0.0, 1e-3],          # hydrostatic strain
    [0.0, 1.0],           # stacking-fault shift fraction
    [-0.05, 0.05],        # rocking angle [deg]
])


def sample_params(n, rng):
    u = rng.random((n, N_IN))
    return RANGES[:, 0] + u * (RANGES[:, 1] - RANGES[:, 0])


def normalise(params):
    return (2.0 * (np.asarray(params) - RANGES[:, 0]) / (RANGES[:, 1] - RANGES[:, 0]) - 1.0
            ).astype(np.float32)


def features(params):
    """(B, 9) defect parameters -> (B, 11) float32 network inputs in [-1, 1]:
    core x, y, Burgers unit vector, line unit vector, strain, fault shift, angle."""
    P = np.atleast_2d(np.asarray(params, dtype=np.float64))
    z = normalise(P).astype(np.float64)
    out = np.empty((P.shape[0], N_FEAT))
    out[:, 0:2] = z[:, 0:2]
    for c, (t, f) in ((2, (2, 3)), (5, (4, 5))):
        th, ph = P[:, t], P[:, f]
        out[:, c] = np.sin(th) * np.cos(ph)
        out[:, c + 1] = np.sin(th) * np.sin(ph)
        out[:, c + 2] = np.cos(th)
    out[:, 8:11] = z[:, 6:9]
    return out.astype(np.float32)


def _unit(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi),
                     math.cos(theta)])


def defect_args(p, orient):
    """The 32-float argument block of mosaic_defect_positions for one draw.

    The fault plane sits a quarter (111) spacing off the origin, between atomic
    planes of the centred lattice, so no atom lies on it."""
    line = _unit(p[4], p[5])
    b = (A_AU / math.sqrt(2.0)) * _unit(p[2], p[3])
    bs = float(b @ line)
    be_vec = b - bs * line
    be = float(np.linalg.norm(be_vec))
    if be > 1e-9:
        e1 = be_vec / be
    else:
        t = np.array([1.0, 0.0, 0.0]) if abs(line[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = t - (t @ line) * line
        e1 /= np.linalg.norm(e1)
    e2 = np.cross(line, e1)
    n = np.array([1.0, 1.0, 1.0]) / math.sqrt(3.0)
    shift = p[7] * (A_AU / 6.0) * np.array([1.0, 1.0, -2.0])
    core = np.array([p[0] * 10.0, p[1] * 10.0, 0.0])
    arr = np.concatenate([core, line, e1, e2, [bs, be, NU_AU, p[6]], n, shift,
                          np.asarray(orient, dtype=np.float64).reshape(9),
                          [0.25 * A_AU / math.sqrt(3.0)]])
    return np.ascontiguousarray(arr, dtype=np.float32)


def defect_positions(base, p, orient, out=None):
    """Displaced, oriented positions (N, 3) float32 on the device."""
    if out is None:
        out = torch.empty_like(base)
    a = defect_args(p, orient)
    s = torch.cuda.current_stream()
    check(_lib.lib().mosaic_defect_positions(int(base.shape[0]), ctypes.c_void_p(base.data_ptr()),
                                             a.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_void_p(out.data_ptr()),
                                             ctypes.c_void_p(int(s.cuda_stream))),
          "mosaic_defect_positions")
    return out


class ImageGenerator:
    """Kinematic images of one defected FCC Au crystal family at (111)."""

    def __init__(self, cells=(37, 37, 37), npix=128, device=None):
        from . import workloads as W
        from .sample import FCC_BASIS, lattice, orient_for_reflection, bragg_angle
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        w = W.fcc_fault_scan111(cells=cells, npix=npix, n_angles=1, fault=False)
        self.w = w
        lam = w["beam"].wavelength
        th = bragg_angle(A_AU / math.sqrt(3.0), lam)
        base, _ = lattice(A_AU, cells, FCC_BASIS)
        self.orient = orient_for_reflection((1, 1, 1), th)
        self.base = torch.as_tensor(base.astype(np.float32), device=self.device)
        self.n = base.shape[0]
        self.codes = torch.zeros(self.n, dtype=torch.uint8, device=self.device)
        self.counts = np.array([self.n], dtype=np.int64)
        b, d = w["beam"], w["detector"]
        self.k = 2.0 * np.pi / lam
        self.Ny, self.Nz = d.shape
        self.spc = K.species_table(tables.species_constants(["Au"], b.energy))
        self.opts = K.kin_opts(self.k, False, True, b._polarization_state(), False)
        self.coords = torch.as_tensor(np.ascontiguousarray(d.pixel_coordinates / 1e10),
                                      device=self.device)
        self.stage = w["stage"]
        self._staged = {}
        self._poses = {}
        self._pos = torch.empty_like(self.base)

    def pose(self, angle_deg):
        """(R, T) of the stage at a rocking angle, cached per angle (the stage
        model costs milliseconds per pose on the host)."""
        key = float(angle_deg)
        got = self._poses.get(key)
        if got is None:
            self.stage.set_single_motor_value_absolute("phi", key)
            got = (np.asarray(self.stage.rotation, np.float32),
                   np.asarray(self.stage.translation, np.float32))
            self.stage.set_single_motor_value_absolute("phi", 0.0)
            if len(self._poses) < 65536:
                self._poses[key] = got
        return got

    def stage_crystal(self, p):
        defect_positions(self.base, p, self.orient, out=self._pos)
        return self._pos

    def intensities(self, p, angles_deg):
        """|A|^2 (F, Nz*Ny) float32 for defect parameters p at the given angles."""
        pos = self.stage_crystal(p)
        F = len(angles_deg)
        poses = [self.pose(a) for a in angles_deg]
        rot = np.stack([q[0] for q in poses])
        tr = np.stack([q[1] for q in poses])
        staged = self._staged.get(F)
        if staged is None:
            # k-d tiles of the undisplaced lattice, kept for every defect: the
            # displacements (a few A) only widen the tiles slightly
            staged = K.stage(self.base, self.codes, None, self.counts, self.k, F, self.Ny,
                             self.Nz)
            self._staged[F] = staged
        K.restage(staged, pos)
        out = torch.zeros((F, self.Nz, self.Ny), dtype=torch.complex64, device=self.device)
        K.kin_sum(staged, self.spc, self.opts, rot, tr, self.coords, out)
        self.path_used = staged.path_used
        return (out.abs() ** 2).reshape(F, -1)

    def probe(self, p, angles_deg, pix):
        """Exact |A|^2 (F, P) at detector pixels ``pix`` (P flat indices,
        device) for defect parameters p at the given angles: the plain sum's
        probe form, one lane per (angle, pixel)."""
        pos = self.stage_crystal(p)
        F, P = len(angles_deg), int(pix.numel())
        poses = [self.pose(a) for a in angles_deg]
        rot = np.stack([q[0] for q in poses])
        tr = np.stack([q[1] for q in poses])
        key = ("probe", F, P)
        staged = self._staged.get(key)
        if staged is None:
            staged = K.stage(self.base, self.codes, None, self.counts, self.k, F, P, 1)
            self._staged[key] = staged
            self._probe_opts = K.kin_opts(self.k, False, True,
                                          self.w["beam"]._polarization_state(), False,
                                          path="probe")
        K.restage(staged, pos)
        coords = self.coords[:, pix].contiguous()
        out = torch.zeros((F, 1, P), dtype=torch.complex64, device=self.device)
        K.kin_sum(staged, self.spc, self._probe_opts, rot, tr, coords, out)
        return (out.abs() ** 2).reshape(F, P)


def make_training_set(gen, n, seed, report_every=2000):
    """n (params, intensity) pairs from the GPU kinematic sum."""
    rng = np.random.default_rng(seed)
    P = sample_params(n, rng)
    I = torch.empty((n, gen.Ny * gen.Nz), dtype=torch.float32, device=gen.device)
    t0 = time.perf_counter()
    for i in range(n):
        I[i] = gen.intensities(P[i], [P[i, 8]])[0]
        if report_every and (i + 1) % report_every == 0:
            torch.cuda.synchronize()
            print(f"[surrogate] generated {i + 1}/{n} images "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
    return P, I


class Layer:
    def __init__(self, E, n_in, n_out, bias_n, act, rng, device):
        std = math.sqrt(2.0 / n_in) if act else math.sqrt(1.0 / n_in)
        self.W = torch.as_tensor(rng.standard_normal((E, n_in, n_out)).astype(np.float32) * std,
                                 device=device)
        self.b = torch.zeros((E, bias_n), dtype=torch.float32, device=device)
        self.act = act
        self.params = [self.W, self.b]
        self.grads = [torch.zeros_like(self.W), torch.zeros_like(self.b)]
        self.m = [torch.zeros_like(self.W), torch.zeros_like(self.b)]
        self.v = [torch.zeros_like(self.W), torch.zeros_like(self.b)]


class SurrogateEnsemble:
    """Ensemble MLP + transposed-conv decoder, all on the surrogate kernels."""

    MLP = (N_FEAT, 256, 512, 8 * 8 * 64)
    CHANNELS = (64, 64, 32, 16, 16, 16)

    def __init__(self, members=5, seed=0, device=None, out_hw=128, target="log"):
        if out_hw < 16 or out_hw & (out_hw - 1):
            raise ValueError("out_hw must be a power of two >= 16")
        if target not in ("rel", "log"):
            raise ValueError("target is 'rel' or 'log'")
        self.target = target
        self.out_hw = out_hw
        n_stage = int(round(math.log2(out_hw // 8)))
        # decoder stages (H, Cin, Cout): 8x8x64 doubling up to out_hw, one channel out
        self.DEC = tuple((8 << i, self.CHANNELS[i],
                          1 if i == n_stage - 1 else self.CHANNELS[i + 1]) for i in range(n_stage))
        self.E = members
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        rng = np.random.default_rng(seed)
        self.mlp = [Layer(members, a, b, b, 1, rng, self.device)
                    for a, b in zip(self.MLP[:-1], self.MLP[1:])]
        self.dec = [Layer(members, ci, 4 * co, co, int(i < len(self.DEC) - 1), rng, self.device)
                    for i, (h, ci, co) in enumerate(self.DEC)]
        self.step_count = 0
        self.norm = None     # (I0, mu, sd)
        self._packed = None  # (step_count, packed weights) for the fused inference
        self._infer_ws = None

    FUSED_DEC = ((8, 64, 64), (16, 64, 32), (32, 32, 16), (64, 16, 1))

    def fused(self):
        """Whether the fused inference kernels cover this architecture."""
        return self.MLP == (N_FEAT, 256, 512, 4096) and self.DEC == self.FUSED_DEC

    def infer(self, X, intensity=False):
        """X (B, 11) ``features`` on the device -> (E, B, 128*128) outputs in
        two launches (mosaic_sg_infer), or with ``intensity`` (log target)
        the intensities, the target's inverse fused into the store; the
        weights are packed into the matrix cores' operand order once per
        weight update."""
        if not self.fused():
            raise ValueError("the fused inference covers the 128 x 128 network only")
        if self._packed is None or self._packed[0] != self.step_count:
            packed = [S.pack(L.W) for L in self.layers]
            self._packed = (self.step_count, packed,
                            S.infer_net(packed, [L.b for L in self.layers]))
        B = int(X.shape[0])
        nb = S.infer_workspace_bytes(self.E, B)
        if self._infer_ws is None or self._infer_ws.numel() < nb:
            self._infer_ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        out = torch.empty((self.E, B, self.out_hw * self.out_hw), dtype=torch.float32,
                          device=self.device)
        post = None
        if intensity:
            I0, mu, sd = self.norm
            post = (sd, mu, I0, -I0 * LOG_EPS)
        return S.infer(X.contiguous(), self._packed[2], out, self._infer_ws, post)

    @property
    def layers(self):
        return self.mlp + self.dec

    def n_params(self):
        """Parameters of one member."""
        return sum(int(p[0].numel()) for L in self.layers for p in L.params)

    # ------------------------------------------------------------ forward
    def forward(self, X):
        """X (B, 11) ``features`` on the device -> (E, B, 128*128) outputs."""
        B = X.shape[0]
        E = self.E
        self._X = X.reshape(1, B, N_FEAT).contiguous()
        self._acts = []
        h = self._X
        for L in self.mlp:
            n_out = L.W.shape[2]
            Z = torch.empty((E, B, n_out), dtype=torch.float32, device=self.device)
            S.gemm(h, L.W, Z)
            Y = torch.empty_like(Z)
            S.bias_act(Z, Y, L.b, n_out, L.act)
            self._acts.append((h, Z))
            h = Y
        img = h.reshape(E, B * 64, 64)
        for i, (L, (H, ci, co)) in enumerate(zip(self.dec, self.DEC)):
            G = torch.empty((E, B * H * H, 4 * co), dtype=torch.float32, device=self.device)
            S.gemm(img, L.W, G)
            Ya = torch.empty_like(G)
            S.bias_act(G, Ya, L.b, co, L.act)
            out = torch.empty((E, B * 4 * H * H, co), dtype=torch.float32, device=self.device)
            S.shuffle2(Ya, out, B, H, H, co)
            self._acts.append((img, G))
            img = out
        return img.reshape(E, B, -1)

    # ----------------------------------------------------------- backward
    def backward(self, dP):
        """dP (E, B, 128*128) -> parameter gradients (summed over the batch)."""
        E = self.E
        B = dP.shape[1]
        d = dP.reshape(E, -1, 1).contiguous()
        for i in reversed(range(len(self.dec))):
            L = self.dec[i]
            H, ci, co = self.DEC[i]
            img, G = self._acts[len(self.mlp) + i]
            dYa = torch.empty_like(G)
            S.shuffle2(dYa, d.reshape(E, -1).contiguous(), B, H, H, co, backward=True)
            dG = torch.empty_like(G)
            S.act_bwd(G, dYa, dG, L.act)
            S.bias_grad(dG, L.grads[1], co)
            S.gemm(img, dG, L.grads[0], transA=True)
            d = torch.empty_like(img)
            S.gemm(dG, L.W, d, transB=True)
        d = d.reshape(E, B, -1).contiguous()
        for i in reversed(range(len(self.mlp))):
            L = self.mlp[i]
            h, Z = self._acts[i]
            dZ = torch.empty_like(Z)
            S.act_bwd(Z, d, dZ, L.act)
            S.bias_grad(dZ, L.grads[1], Z.shape[2])
            if h.shape[0] == 1 and E > 1:
                hb = h.expand(E, -1, -1).contiguous()
            else:
                hb = h
            S.gemm(hb, dZ, L.grads[0], transA=True)
            if i > 0:
                d = torch.empty_like(h)
                S.gemm(dZ, L.W, d, transB=True)

    def adam_step(self, lr, weight_decay=0.0):
        self.step_count += 1
        for L in self.layers:
            for p, g, m, v in zip(L.params, L.grads, L.m, L.v):
                S.adam(p, g, m, v, lr, self.step_count, weight_decay=weight_decay)

    # ------------------------------------------------------------- fitting
    def set_normalisation(self, I):
        I0 = float(I.max().item())
        if self.target == "rel":
            self.norm = (I0, 0.0, 1.0)
            return
        y = torch.log(I / I0 + LOG_EPS)
        self.norm = (I0, float(y.mean().item()), float(y.std().item()))

    def targets(self, I):
        I0, mu, sd = self.norm
        if self.target == "rel":
            return (I / I0).contiguous()
        return ((torch.log(I / I0 + LOG_EPS) - mu) / sd).contiguous()

    def intensity(self, y):
        I0, mu, sd = self.norm
        if self.target == "rel":
            return I0 * torch.clamp(y, min=0.0)
        return I0 * (torch.exp(y * sd + mu) - LOG_EPS)

    def train(self, P, I, epochs, batch=128, lr=2e-3, seed=0, log=print, lr_min=1e-5):
        """Adam on each image's squared relative L2 error (``target="rel"``)
        or on the mean squared error of the normalised log-intensity."""
        if self.norm is None:
            self.set_normalisation(I)
        n = P.shape[0]
        X = torch.as_tensor(features(P), device=self.device)
        Y = self.targets(I)
        # per-image weights 1 / mean(y^2): the loss is the squared relative error
        Wt = (1.0 / (Y * Y).mean(dim=1).clamp_min(1e-30)).contiguous() \
            if self.target == "rel" else None
        rng = np.random.default_rng(seed)
        loss = torch.zeros(self.E, dtype=torch.float32, device=self.device)
        steps = epochs * (n // batch)
        hist = []
        s = 0
        for ep in range(epochs):
            perm = torch.as_tensor(rng.permutation(n), device=self.device)
            loss.zero_()
            nb = 0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                out = self.forward(X[idx])
                dP = torch.empty_like(out)
                # loss and gradient of the mean over the batch's pixels
                S.mse(out, Y[idx].reshape(-1).contiguous(), dP, loss,
                      row_weight=Wt[idx].contiguous() if Wt is not None else None)
                self.backward(dP)
                cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * s / max(1, steps)))
                self.adam_step(cur)
                s += 1
                nb += 1
            ep_loss = (loss / nb).cpu().numpy()
            hist.append(ep_loss)
            if log:
                log(f"[surrogate] epoch {ep + 1}/{epochs} loss per member "
                    f"{np.array2string(ep_loss, precision=4)}")
        return np.array(hist)

    def predict(self, P, chunk=256):
        """Ensemble mean intensity (B, npix) and per-frame spread (B,), in
        chunks of ``chunk`` frames (the decoder GEMMs have B x H x W rows).
        The spread is the members' RMS deviation over the frame, relative to
        the RMS of the mean with ``target="rel"`` (comparable with the
        relative L2 error)."""
        Xall = torch.as_tensor(features(P), device=self.device)
        means, spreads = [], []
        run = self.infer if self.fused() else self.forward
        for i in range(0, Xall.shape[0], chunk):
            out = run(Xall[i:i + chunk].contiguous())
            B = out.shape[1]
            mean = torch.empty((B, out.shape[2]), dtype=torch.float32, device=self.device)
            spread = torch.empty(B, dtype=torch.float32, device=self.device)
            S.ensemble_stats(out.contiguous(), mean, spread)
            if self.target == "rel":
                spread = spread / (mean * mean).mean(dim=1).sqrt().clamp_min(1e-30)
            means.append(self.intensity(mean))
            spreads.append(spread)
        return torch.cat(means), torch.cat(spreads)


    def member_intensities(self, P, chunk=256):
        """Each member's intensity (E, B, npix) for (B, 9) defect parameters."""
        Xall = torch.as_tensor(features(P), device=self.device)
        outs = []
        fused = self.fused() and self.target == "log"
        for i in range(0, Xall.shape[0], chunk):
            x = Xall[i:i + chunk].contiguous()
            outs.append(self.infer(x, intensity=True) if fused
                        else self.intensity(self.infer(x) if self.fused() else self.forward(x)))
        return torch.cat(outs, dim=1)


def rel_l2(a, b):
    """Relative L2 error along the last axis, in float64 (intensities in r_e
    units are ~1e-18: their squares underflow float32)."""
    a, b = a.double(), b.double()
    return (torch.linalg.norm(a - b, dim=-1) / torch.linalg.norm(b, dim=-1)).float()


def anchor_weights(angles, anchor_angles):
    """(F, A) weights interpolating anchor values linearly in angle (held
    constant beyond the outer anchors)."""
    a = np.asarray(angles, dtype=np.float64)
    x = np.asarray(anchor_angles, dtype=np.float64)
    W = np.zeros((a.size, x.size))
    for f, t in enumerate(a):
        if x.size == 1 or t <= x[0]:
            W[f, 0] = 1.0
        elif t >= x[-1]:
            W[f, -1] = 1.0
        else:
            j = int(np.searchsorted(x, t, side="right")) - 1
            u = (t - x[j]) / (x[j + 1] - x[j])
            W[f, j], W[f, j + 1] = 1.0 - u, u
    return W.astype(np.float32)


class SurrogateScan:
    """A rocking scan from the ensemble, anchored on plain frames and judged
    by exact probes.

    The plain sum computes the anchor frames (by default the first and the
    last angle).  Each member's prediction is corrected by that member's own
    error at the anchors, interpolated linearly in angle, and the corrected
    members' mean is the frame.  The judge then evaluates the exact field at
    ``judge_pixels`` detector pixels of every other frame (the probe form of
    the plain sum), drawn with replacement with probability q_p proportional
    to the predicted scan-mean intensity plus 5 % of its mean, and estimates
    each frame's relative L2 error by the importance-weighted ratio

        e_f^2 = sum_k (I'_fk - I_fk)^2 / q_k / sum_k I_fk^2 / q_k.

    A frame is kept when the members agree (their RMS spread about the
    frame, relative to its RMS, at most ``spread_limit``) and the estimate's
    upper bound, e_f^2 plus ``sigmas`` standard errors of the
    importance-sampled ratio (delta method over the draws), stays within
    ``threshold`` squared; the others are recomputed with the plain sum.  The anchors also judge the network: when its raw error at an anchor
    exceeds ``anchor_limit`` the whole scan is computed with the plain sum.
    """

    def __init__(self, model, gen, threshold=1e-2, anchors=(0, -1), anchor_limit=0.25,
                 judge_pixels=64, seed=0, sigmas=3.0, spread_limit=0.05):
        self.model, self.gen = model, gen
        self.threshold = float(threshold)
        self.spread_limit = float(spread_limit)
        self.last_spread = None
        self.sigmas = float(sigmas)
        self.anchors = tuple(anchors)
        self.anchor_limit = float(anchor_limit)
        self.judge_pixels = int(judge_pixels)
        self._gen = torch.Generator(device=gen.device)
        self._gen.manual_seed(seed)

    def run(self, p, angles_deg):
        """(I (F, npix), estimated relative error per frame (F,; 0 for plain
        frames), indices recomputed with the plain sum).  The host waits on
        the device once, for the anchors' verdict and the rejected frames."""
        angles = np.asarray(angles_deg, dtype=np.float64)
        F = angles.size
        ia = sorted({a % F for a in self.anchors})
        inner = np.array([f for f in range(F) if f not in ia], dtype=np.int64)
        P = np.repeat(np.asarray(p, dtype=np.float64)[None], F, axis=0)
        P[:, 8] = angles
        Im = self.model.member_intensities(P)                    # (E, F, npix)
        Ia = self.gen.intensities(p, angles[ia])                 # (A, npix), plain
        dev = Ia.device
        ia_t = torch.as_tensor(ia, dtype=torch.int64, device=dev)
        in_t = torch.as_tensor(inner, dtype=torch.int64, device=dev)
        raw = rel_l2(Im[:, ia_t].mean(dim=0), Ia)
        W = torch.as_tensor(anchor_weights(angles, angles[ia]), device=dev)
        C = Im + torch.matmul(W, Ia[None] - Im[:, ia_t])         # corrected members
        mean = C.mean(dim=0)
        # the members' disagreement, relative to the frame, on intensities
        # scaled by the anchors' peak (in r_e units they are ~1e-18, whose
        # squares float32 cannot hold to full precision)
        inv = 1.0 / Ia.max().clamp_min(1e-37)
        Cs, ms = C * inv, mean * inv
        spread = torch.sqrt((Cs - ms[None]).pow(2).mean(dim=(0, 2)) /
                            ms.pow(2).mean(dim=1).clamp_min(1e-30))
        spread[ia_t] = 0.0
        self.last_spread = spread
        mean = mean.clamp_min(0.0)
        mean[ia_t] = Ia
        est = torch.zeros(F, dtype=torch.float32, device=dev)
        upper = torch.zeros_like(est)
        if inner.size and self.judge_pixels > 0:
            mi = mean[in_t]
            w = mi.mean(dim=0)
            w = w + 0.05 * w.mean() + 1e-30
            q = w / w.sum()
            pix = torch.multinomial(q, self.judge_pixels, replacement=True,
                                    generator=self._gen)
            # float64: intensities in r_e units are ~1e-18, their squares
            # and squared differences underflow float32
            exact = self.gen.probe(p, angles[inner], pix).double()     # (Fi, P)
            iw = 1.0 / q[pix].double()
            d = mi[:, pix].double() - exact
            a_ = d * d * iw                       # per-draw terms of the ratio
            b_ = exact * exact * iw
            ma, mb = a_.mean(dim=1), b_.mean(dim=1)
            mb = torch.where(mb > 0, mb, torch.ones_like(mb))
            R = ma / mb
            # delta-method variance of the ratio of means over the n draws
            n = a_.shape[1]
            va = ((a_ - ma[:, None]) - R[:, None] * (b_ - mb[:, None])).pow(2).sum(dim=1)
            var = va / (max(n - 1, 1) * n) / (mb * mb)
            est[in_t] = torch.sqrt(R).float()
            upper[in_t] = torch.sqrt(R + self.sigmas * torch.sqrt(var)).float()
        flags = torch.cat([(raw.max() > self.anchor_limit).reshape(1),
                           (upper > self.threshold) | (spread > self.spread_limit)]).cpu().numpy()
        if flags[0]:
            est = torch.full((F,), float("inf"), device=dev)
            return self.gen.intensities(p, angles), est, np.arange(F)
        bad = np.nonzero(flags[1:])[0]
        if bad.size:
            mean[torch.as_tensor(bad, device=dev)] = self.gen.intensities(p, angles[bad])
        return mean, est, bad
