"""Surrogate kernels against plain PyTorch float32 references.

The GEMM rounds its operands to bfloat16 and accumulates in float32, so it is
compared with torch's float32 product of the bf16-rounded operands
(tolerance 1e-5 relative) and, loosely, with the unrounded product.
"""
import numpy as np
import pytest
import torch

from mosaic_amd import sg_ops as S

pytestmark = pytest.mark.gpu


def bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("shape", [(3, 70, 45, 37), (1, 129, 200, 64), (2, 1, 5, 300)])
def test_gemm_vs_torch(device, ta, tb, shape):
    nb, M, N, K = shape
    g = torch.Generator(device="cpu").manual_seed(M * 7 + N)
    A = torch.randn((nb, K, M) if ta else (nb, M, K), generator=g).to(device)
    B = torch.randn((nb, N, K) if tb else (nb, K, N), generator=g).to(device)
    C = torch.randn((nb, M, N), generator=g).to(device)
    C0 = C.clone()
    S.gemm(A, B, C, transA=bool(ta), transB=bool(tb), alpha=0.5, beta=0.25)
    Ao = bf16(A).transpose(1, 2) if ta else bf16(A)
    Bo = bf16(B).transpose(1, 2) if tb else bf16(B)
    ref = 0.5 * torch.matmul(Ao.contiguous(), Bo.contiguous()) + 0.25 * C0
    err = (C - ref).abs().max().item() / ref.abs().max().item()
    print(f"[parity] gemm {shape} ta={ta} tb={tb}: rel={err:.2e}")
    assert err < 1e-5


def test_bias_act_and_backward(device):
    Z = torch.randn(2, 33, 12, device=device)
    bias = torch.randn(2, 4, device=device)
    Zc = Z.clone()
    Y = torch.empty_like(Z)
    S.bias_act(Z, Y, bias, 4, 1)
    ref_z = Zc + bias.repeat(1, 3)[:, None, :]
    assert torch.allclose(Z, ref_z, atol=1e-6)
    assert torch.allclose(Y, torch.nn.functional.silu(ref_z), atol=1e-5)
    dY = torch.randn_like(Z)
    dZ = torch.empty_like(Z)
    S.act_bwd(Z, dY, dZ, 1)
    zz = ref_z.clone().requires_grad_(True)
    torch.nn.functional.silu(zz).backward(dY)
    assert torch.allclose(dZ, zz.grad, atol=1e-5)
    db = torch.empty(2, 4, device=device)
    S.bias_grad(dZ, db, 4)
    assert torch.allclose(db, dZ.reshape(2, 33, 3, 4).sum(dim=(1, 2)), rtol=1e-5, atol=1e-5)


def test_shuffle_roundtrip(device):
    n, H, W, Co = 3, 4, 5, 6
    G = torch.randn(2, n * H * W, 4 * Co, device=device)
    Y = torch.empty(2, n * 2 * H * 2 * W * Co, device=device)
    S.shuffle2(G, Y, n, H, W, Co)
    ref = G.reshape(2, n, H, W, 2, 2, Co).permute(0, 1, 2, 4, 3, 5, 6).reshape(2, -1)
    assert torch.equal(Y, ref)
    G2 = torch.empty_like(G)
    S.shuffle2(G2, Y, n, H, W, Co, backward=True)
    assert torch.equal(G2, G)


def test_mse_adam_stats(device):
    P = torch.randn(3, 50, device=device)
    T = torch.randn(50, device=device)
    dP = torch.empty_like(P)
    loss = torch.zeros(3, device=device)
    S.mse(P, T, dP, loss)
    assert torch.allclose(loss, ((P - T) ** 2).mean(dim=1), rtol=1e-5)
    assert torch.allclose(dP, 2 * (P - T) / 50, atol=1e-6)
    # row weights: T as 5 rows of 10
    rw = torch.rand(5, device=device) + 0.5
    loss.zero_()
    S.mse(P, T, dP, loss, row_weight=rw)
    wfull = rw.repeat_interleave(10)
    assert torch.allclose(loss, (wfull * (P - T) ** 2).mean(dim=1), rtol=1e-5)
    assert torch.allclose(dP, 2 * wfull * (P - T) / 50, atol=1e-6)
    w = torch.randn(100, device=device)
    g = torch.randn(100, device=device)
    m = torch.zeros(100, device=device)
    v = torch.zeros(100, device=device)
    w_ref = w.clone().requires_grad_(True)
    opt = torch.optim.Adam([w_ref], lr=1e-2)
    for step in range(1, 4):
        S.adam(w, g, m, v, 1e-2, step)
        w_ref.grad = g.clone()
        opt.step()
    assert torch.allclose(w, w_ref.detach(), atol=1e-6)
    Pm = torch.randn(5, 4, 30, device=device)
    mean = torch.empty(4, 30, device=device)
    spread = torch.empty(4, device=device)
    S.ensemble_stats(Pm, mean, spread)
    assert torch.allclose(mean, Pm.mean(0), atol=1e-6)
    assert torch.allclose(spread, Pm.var(0, unbiased=False).mean(1).sqrt(), atol=1e-5)


@pytest.mark.parametrize("rows,cols,period", [(300000, 4, 1), (5000, 256, 64), (7, 12, 4)])
def test_bias_grad_split_rows(device, rows, cols, period):
    """The bias gradient split over workgroups (atomic partial sums) against
    torch's float64 sum, after a stale value in the output."""
    g = torch.Generator(device="cpu").manual_seed(rows)
    dZ = torch.randn((3, rows, cols), generator=g).to(device)
    db = torch.full((3, period), 7.0, device=device)
    S.bias_grad(dZ, db, period)
    ref = dZ.double().reshape(3, rows, cols // period, period).sum(dim=(1, 2))
    err = ((db.double() - ref).abs().max() / ref.abs().max()).item()
    print(f"[parity] bias_grad rows={rows} cols={cols} period={period}: rel={err:.2e}")
    assert err < 1e-5
