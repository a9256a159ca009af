"""The allocation-free entry point mosaic_scatter_noalloc against
mosaic_scatter, through the C ABI: two calls in a row on the same GPU, the
second with a smaller input (the scratch keeps the first call's contents),
plus the size limit.  Bar: the float32 outputs agree to 1e-6 relative (the
same kernels; only the number of atom slices may differ).
"""
import ctypes

import numpy as np
import pytest
import torch

from mosaic_amd import _lib
from mosaic_amd import kinematic as K
from mosaic_amd import workloads as W

pytestmark = pytest.mark.gpu


def _inputs(w, device):
    a = W.oracle_inputs(w)
    dev = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(device)
    t = dict(pos_m=dev(a["pos_m"], np.float64), f0p=dev(a["f0p"], np.float32),
             f0z=dev(a["f0z"], np.float32), sanr=dev(a["sanr"], np.float32),
             sani=dev(a["sani"], np.float32), amp_r=dev(a["amp_r"], np.float32),
             amp_i=dev(a["amp_i"], np.float32), cx=dev(a["cx"], np.float64),
             cy=dev(a["cy"], np.float64), cz=dev(a["cz"], np.float64))
    return t, a


def _noalloc(t, a, flags, out_r, out_i):
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    rc = _lib.lib().mosaic_scatter_noalloc(
        int(t["pos_m"].shape[0]), p(t["pos_m"]), p(t["f0p"]), p(t["f0z"]),
        int(flags["remove_forward"]), p(t["sanr"]), p(t["sani"]), p(t["amp_r"]),
        p(t["amp_i"]), int(a["Ny"]), int(a["Nz"]), p(t["cx"]), p(t["cy"]), p(t["cz"]),
        float(a["k_val"]), int(flags["apply_pol"]), float(a["pol_deg"]), float(a["pol_e"][0]),
        float(a["pol_e"][1]), float(a["pol_e"][2]), int(flags["apply_decay"]), p(out_r),
        p(out_i), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


def _with_workspace(t, a, flags, out_r, out_i):
    K.scatter(t["pos_m"], t["f0p"], t["f0z"], t["sanr"], t["sani"], t["amp_r"], t["amp_i"],
              t["cx"], t["cy"], t["cz"], a["k_val"], flags["remove_forward"], a["Ny"], a["Nz"],
              flags["apply_pol"], a["pol_deg"], a["pol_e"], flags["apply_decay"], out_r, out_i)


def _field(r, i):
    return r.cpu().numpy().astype(np.float64) + 1j * i.cpu().numpy()


def test_noalloc_matches_workspace_entry_twice(device):
    flags = dict(remove_forward=True, apply_pol=True, apply_decay=True)
    cases = [W.fcc_au_bragg111(n=9, npix=32, seed=1), W.fcc_au_bragg111(n=5, npix=12, seed=2)]
    for w in cases:
        t, a = _inputs(w, device)
        npix = a["Ny"] * a["Nz"]
        print(f"[noalloc] {t['pos_m'].shape[0]} atoms x {npix} pixels", flush=True)
        r0 = torch.zeros(npix, dtype=torch.float32, device=device)
        i0 = torch.zeros_like(r0)
        r1, i1 = torch.zeros_like(r0), torch.zeros_like(r0)
        _with_workspace(t, a, flags, r0, i0)
        assert _noalloc(t, a, flags, r1, i1) == _lib.MOSAIC_OK
        torch.cuda.synchronize()
        want, got = _field(r0, i0), _field(r1, i1)
        e = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        print(f"[noalloc] rel {e:.2e}", flush=True)
        assert np.linalg.norm(want) > 0 and e <= 1e-6


def test_noalloc_rejects_oversized_detector(device):
    w = W.fcc_au_bragg111(n=2, npix=4, seed=0)
    t, a = _inputs(w, device)
    out = torch.zeros(16, dtype=torch.float32, device=device)
    a_big = dict(a, Ny=2049, Nz=2048)          # 16 B x 4.2M pixels > 64 MiB
    assert _noalloc(t, a_big, dict(remove_forward=False, apply_pol=False, apply_decay=False),
                    out, out) == _lib.MOSAIC_ERR_TOO_LARGE
    torch.cuda.synchronize()
    assert not torch.any(out != 0)
