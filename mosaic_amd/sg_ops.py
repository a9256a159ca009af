"""Thin torch-tensor wrappers of the surrogate kernels (mosaic_sg_* in the C ABI).

Every function takes float32 CUDA tensors, checks shapes, and launches on the
current stream.  ``batch`` dimensions are leading ensemble members.
"""
import ctypes

import torch

from . import _lib
from ._lib import check


def _p(t):
    return ctypes.c_void_p(int(t.data_ptr()))


def _s():
    return ctypes.c_void_p(int(torch.cuda.current_stream().cuda_stream))


def _f32(*ts):
    for t in ts:
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("surrogate kernels take contiguous float32 device tensors")


def gemm(A, B, C, transA=False, transB=False, alpha=1.0, beta=0.0):
    """C[b] = alpha op(A[b]) op(B[b]) + beta C[b] for 3-D (batch, rows, cols) tensors."""
    _f32(A, B, C)
    nb = C.shape[0]
    M, N = C.shape[1], C.shape[2]
    K = A.shape[1] if transA else A.shape[2]
    if (A.shape[2] if transA else A.shape[1]) != M or (B.shape[2] if transB else B.shape[1]) != K \
            or (B.shape[1] if transB else B.shape[2]) != N:
        raise ValueError(f"gemm shapes {tuple(A.shape)} {tuple(B.shape)} {tuple(C.shape)}")
    sA = A.shape[1] * A.shape[2] if A.shape[0] > 1 else 0
    sB = B.shape[1] * B.shape[2] if B.shape[0] > 1 else 0
    check(_lib.lib().mosaic_sg_gemm(int(transA), int(transB), M, N, K, _p(A), A.shape[2], sA,
                                    _p(B), B.shape[2], sB, _p(C), N, M * N, nb, float(alpha),
                                    float(beta), _s()), "mosaic_sg_gemm")
    return C


def bias_act(Z, Y, bias, period, act):
    """Z (batch, rows, cols) += bias (batch, period) broadcast; Y = SiLU(Z) or Z."""
    _f32(Z, Y, bias)
    nb, rows, cols = Z.shape
    check(_lib.lib().mosaic_sg_bias_act(_p(Z), _p(Y), _p(bias), rows, cols, int(period),
                                        rows * cols, bias.shape[1], nb, int(act), _s()),
          "mosaic_sg_bias_act")


def act_bwd(Z, dY, dZ, act):
    _f32(Z, dY, dZ)
    check(_lib.lib().mosaic_sg_act_bwd(_p(Z), _p(dY), _p(dZ), Z.numel(), int(act), _s()),
          "mosaic_sg_act_bwd")


def bias_grad(dZ, dbias, period):
    _f32(dZ, dbias)
    nb, rows, cols = dZ.shape
    check(_lib.lib().mosaic_sg_bias_grad(_p(dZ), _p(dbias), rows, cols, int(period), rows * cols,
                                         dbias.shape[1], nb, _s()), "mosaic_sg_bias_grad")


def shuffle2(G, Y, nimg, H, W, Co, backward=False):
    _f32(G, Y)
    nb = G.shape[0]
    check(_lib.lib().mosaic_sg_shuffle2(_p(G), _p(Y), int(nimg), int(H), int(W), int(Co),
                                        G[0].numel(), nb, int(backward), _s()),
          "mosaic_sg_shuffle2")


def mse(P, T, dP, loss, row_weight=None):
    """Per member b of P (members, ...): loss[b] += mean(w (P_b - T)^2) and
    dP_b = its gradient; ``row_weight`` (rows,) weighs each row of T's last axis."""
    _f32(P, T, dP, loss)
    nb = P.shape[0]
    n = P[0].numel()
    rl = 0
    if row_weight is not None:
        _f32(row_weight)
        rl = n // row_weight.numel()
        if rl * row_weight.numel() != n:
            raise ValueError("row_weight must have one entry per row of T")
    check(_lib.lib().mosaic_sg_mse(_p(P), _p(T), _p(dP), _p(loss), n, n, nb,
                                   _p(row_weight) if row_weight is not None else None, rl,
                                   _s()), "mosaic_sg_mse")


def adam(w, g, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
    _f32(w, g, m, v)
    check(_lib.lib().mosaic_sg_adam(_p(w), _p(g), _p(m), _p(v), w.numel(), float(lr),
                                    float(beta1), float(beta2), float(eps), int(step),
                                    float(weight_decay), _s()), "mosaic_sg_adam")


def ensemble_stats(P, mean, spread):
    """P (members, frames, npix) -> mean (frames, npix), spread (frames,)."""
    _f32(P, mean, spread)
    members, frames, npix = P.shape
    check(_lib.lib().mosaic_sg_ensemble_stats(_p(P), _p(mean), _p(spread), members, frames, npix,
                                              frames * npix, _s()), "mosaic_sg_ensemble_stats")


def pack(W):
    """(E, K, N) float32 weights -> packed bf16 operand blocks (int16 tensor of
    E x ceil(K/16)*16 x ceil(N/32)*32 elements) for ``infer``."""
    _f32(W)
    E, K, N = W.shape
    Kp, Np = (K + 15) // 16 * 16, (N + 31) // 32 * 32
    out = torch.empty(E * Kp * Np, dtype=torch.int16, device=W.device)
    check(_lib.lib().mosaic_sg_pack(E, K, N, _p(W), K * N, _p(out), _s()), "mosaic_sg_pack")
    return out


def infer_workspace_bytes(members, batch):
    n = ctypes.c_size_t(0)
    check(_lib.lib().mosaic_sg_infer_workspace(members, batch, ctypes.byref(n)),
          "mosaic_sg_infer_workspace")
    return int(n.value)


def infer_net(packed, biases):
    """The mosaic_sg_net struct of ``packed`` (the 7 packed weight tensors)
    and ``biases`` (the 7 (E, n) bias tensors) in layer order; the tensors
    must outlive it."""
    _f32(*biases)
    net = _lib.SgNet()
    for name, t in zip(("W1", "W2", "W3", "D1", "D2", "D3", "D4"), packed):
        setattr(net, name, int(t.data_ptr()))
    for name, t in zip(("b1", "b2", "b3", "d1", "d2", "d3", "d4"), biases):
        setattr(net, name, int(t.data_ptr()))
    return net


def infer(X, net, out, workspace, post=None):
    """Fused forward of every member: X (B, 11), ``net`` from ``infer_net``,
    out (E, B, 128*128); ``post`` = (a, b, c, d) stores c exp(a y + b) + d
    in place of each output y."""
    _f32(X, out)
    E, B = out.shape[0], out.shape[1]
    if X.shape != (B, 11) or out.shape[2] != 128 * 128:
        raise ValueError(f"infer shapes {tuple(X.shape)} {tuple(out.shape)}")
    pm = None
    if post is not None:
        pm = (ctypes.c_float * 4)(*[float(x) for x in post])
    check(_lib.lib().mosaic_sg_infer(E, B, _p(X), ctypes.byref(net),
                                     ctypes.cast(pm, ctypes.c_void_p) if pm is not None else None,
                                     _p(workspace), workspace.numel() * workspace.element_size(),
                                     _p(out), _s()),
          "mosaic_sg_infer")
    return out
