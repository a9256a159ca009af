"""Accuracy of the hardware sine and cosine (v_sin_f32 / v_cos_f32, input in
revolutions) over the argument ranges the sum kernels feed them, against
float64 sin(2 pi x) and cos(2 pi x) of the same float32 argument:
  [0, 1)            the per-tile, recurrence and exact kernels (after fract);
  [-2, 2]           the moment sum's origin phase (host-bounded below 2 rev);
  [-200, 200]       the moment sum's bound on the origin phase range;
  [-256, 256]       the matrix-core operands (2 radius + 1 <= 256, unreduced).
Bar: absolute error <= 2e-6 on each range (the kernels' own FP32 phase
rounding is ~1e-6 rev at 25 rev); the measured figures are printed.
"""
import ctypes

import numpy as np
import pytest
import torch

from mosaic_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lo,hi", [(0.0, 1.0), (-2.0, 2.0), (-200.0, 200.0),
                                   (-256.0, 256.0)])
def test_hardware_sincos_range(device, lo, hi):
    rng = np.random.default_rng(int(hi))
    x = rng.uniform(lo, hi, 1 << 22).astype(np.float32)
    x[:5] = np.float32([lo, np.nextafter(np.float32(hi), np.float32(0)), 0.25, 0.5, -0.75])
    d = torch.from_numpy(x).to(device)
    s = torch.empty_like(d)
    c = torch.empty_like(d)
    p = lambda t: ctypes.c_void_p(int(t.data_ptr()))
    _lib.check(_lib.lib().mosaic_sincos_rev(d.numel(), p(d), p(s), p(c),
                                            ctypes.c_void_p(int(torch.cuda.current_stream()
                                                                .cuda_stream))),
               "mosaic_sincos_rev")
    torch.cuda.synchronize()
    xd = x.astype(np.float64)
    # reduce in float64 exactly: x is a float32, so x - rint(x) is exact
    r = xd - np.rint(xd)
    es = np.abs(s.cpu().numpy() - np.sin(2 * np.pi * r))
    ec = np.abs(c.cpu().numpy() - np.cos(2 * np.pi * r))
    print(f"[sincos] [{lo}, {hi}]: max |err| sin {es.max():.2e} cos {ec.max():.2e}, "
          f"rms {np.sqrt(np.mean(es ** 2)):.2e} / {np.sqrt(np.mean(ec ** 2)):.2e}", flush=True)
    assert es.max() <= 2e-6 and ec.max() <= 2e-6
