"""Benchmark: kinematic sum throughput (atom-pixel pairs/s) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workload (default): the precision-form sample -- SrTiO3-type perovskite,
100^3 cells = 5,000,000 atoms (3 species, 64 spatial chunks) with a screw
dislocation, 10 keV, 512 x 512 pixels at 1 m on the (200) reflection:
1.31e12 atom-pixel pairs per step.  The sample is resident in HBM before
the timed region; one step stages it (k-d cut into tiles), runs the
pixel prologue and the sum.

Prints one JSON line.  ``roofline`` prices the dominant kernel of the sum:
on this workload the matrix-core moment sum's moments kernel (k_gm_moments),
by the FP16 matrix flops it executes per launch (the profile's counters of
the same workload) over its launch time, against the 2.5 PFLOP/s dense FP16
peak, with its FP32 VALU flops and issue fraction beside them (the VALU
kernels are priced by their FP32 flops against the 157.3 TFLOP/s peak).
The launch time comes from the library's HIP events around that kernel
(mosaic_kin_timing) on the stream it runs on; ``sum_ms`` is the whole
mosaic_kin_sum call by HIP events.  ``cpu_baseline`` times the reference's own C kernel
(compute_scattering_cffi, built into oracle/_ref) on a bounded pixel sample
of the same workload, on the host's cores.  ``end_to_end`` times the drop-in
call on the host-resident chunks (H2D included); ``scan``, ``adjoint`` and
``propagation`` cover the other workloads, each with its error against the
oracle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 20.0          # algorithmic FP32 work per atom-pixel pair (20 flop + 2 trans.)
PEAK_FP32 = 157.3e12          # MI355X FP32 vector (packed FMA), the published peak
PEAK_FP16_MFMA = 2.5e15       # MI355X dense FP16 matrix
# VALU issue peak in lane-operations per second (an FMA counts once): 256 CUs x
# 4 SIMDs x 16 lanes x 2.4 GHz, the FP32 peak halved
PEAK_VALU_LANE_OPS = PEAK_FP32 / 2
# v_sin_f32 / v_cos_f32 measured on the box (a probe kernel, 8 chains per lane):
# 1.81e13 lane-operations/s, a quarter of the FMA issue rate
PEAK_TRANS_LANE_OPS = 1.81e13
# The matrix-core sum executes 3 split products x 4 real multiply-adds per pair.
MFMA_FLOP_PER_PAIR = 24.0
# One v_mfma_f32_32x32x16_f16: 32 x 32 x 16 multiply-adds.
MFMA_FLOP_PER_INST = 2.0 * 32 * 32 * 16
MFMA_F32_FLOP_PER_INST = 2.0 * 32 * 32 * 2

# Counters of one launch of the sum kernels on the 5M-atom, 512^2 workload
# (rocprofv3 --pmc, same bench command; the profile files hold the commands):
# FETCH_SIZE / WRITE_SIZE in KiB (scaled by the copy calibration, FETCH x 2,
# WRITE x 1) and the FP32 VALU instruction counts by kind (wave-instructions).
COUNTERS = {
    "k_sum_poly": dict(
        source="profiles/k_sum_poly_perovskite5M_512px.md",
        fetch_kib=719146.0, write_kib=266402.0, valu=4.711e9,
        add_f32=2.06103e8, mul_f32=8.25971e8, fma_f32=2.82295e9, trans_f32=9.04499e7),
    "k_gm_moments": dict(
        source="profiles/k_gm_moments_perovskite5M_512px.md",
        fetch_kib=315430.0, write_kib=440380.0, valu=4.3926e8,
        add_f32=8.3862e6, mul_f32=6.4594e7, fma_f32=4.1348e7, trans_f32=4.0004e7,
        mfma=3.5e7, mfma_f32=5.0e6),
    "k_sum_mma": dict(
        source="profiles/k_sum_mma_perovskite5M_512px.md",
        fetch_kib=1831133.0, write_kib=2542443.8, valu=None,
        add_f32=None, mul_f32=None, fma_f32=None, trans_f32=None),
}
WORKLOAD_OF_COUNTERS = "perovskite_screw_5M_512px"
# The pruned propagation's three passes on the bench's 2048^2-on-8192^2 step:
# FETCH_SIZE, WRITE_SIZE (KiB per launch, same scaling) from
# profiles/propagation_crl2m_2048_on_8192.md.
PROP_COUNTERS = {"k_asp_rows_fwd": (18263.6, 593124.0), "k_asp_cols": (86624.1, 178921.0),
                 "k_asp_rows_inv": (77330.7, 32768.0)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="perovskite_screw_5M_512px")
    ap.add_argument("--cpu-pixels", type=int, default=768)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--err-pixels", type=int, default=48)
    ap.add_argument("--path", default="auto",
                    choices=["auto", "tile", "rows", "mma", "poly", "gemm", "exact"])
    ap.add_argument("--no-propagation", action="store_true")
    ap.add_argument("--no-scan", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-end-to-end", action="store_true")
    ap.add_argument("--no-surrogate-inference", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--surrogate", action="store_true",
                    help="run the surrogate workload (data generation, training, gated scan)")
    ap.add_argument("--sg-images", type=int, default=20000)
    ap.add_argument("--sg-epochs", type=int, default=100)
    ap.add_argument("--sg-scans", type=int, default=8)
    ap.add_argument("--sg-infer-only", action="store_true",
                    help="time only the surrogate's fused inference (5 members x 64 frames)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device 3 times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to HBM bytes)")
    return ap.parse_args()


def build_workload(name):
    from mosaic_amd import workloads as W
    if name == "perovskite_screw_5M_512px":
        return W.perovskite_screw_bragg200(n=100, npix=512)
    if name == "fcc_au_32k_128px":
        return W.fcc_au_bragg111(n=20, npix=128)
    if name.startswith("perovskite_screw_n"):
        n, p = name[len("perovskite_screw_n"):].split("_")
        return W.perovskite_screw_bragg200(n=int(n), npix=int(p))
    raise SystemExit(f"unknown workload {name}")


def gather(w):
    smp = w["sample"]
    labels, index, pos, codes = [], {}, [], []
    for cid in range(1, smp.chunk_total + 1):
        p = smp.load_chunk_positions(cid)
        c, lab = smp.load_chunk_species(cid, codes=True)
        rm = []
        for el in lab:
            el = str(el)
            if el not in index:
                index[el] = len(labels)
                labels.append(el)
            rm.append(index[el])
        pos.append(p)
        codes.append(np.asarray(rm, dtype=np.uint8)[np.asarray(c, dtype=np.int64)])
    return np.concatenate(pos), np.concatenate(codes), labels


def host_cpu():
    """Threads the CPU baseline uses and the host's identity.

    The reference's CPU driver uses every logical core (Beam.py:5873-5955,
    hardware.py:790-803); on a shared host the cores this process may run on
    (its affinity, and OMP_NUM_THREADS where the host sets it) are that set."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity_cpus": affinity,
                     "omp_num_threads": omp, "model": model}


def cpu_baseline(w, n_pix, seed=0):
    from mosaic_amd import workloads as W
    from oracle import kinematic as O
    if not O.ref_available():
        return None
    a = W.oracle_inputs(w)
    rng = np.random.default_rng(seed)
    sel = rng.choice(a["Ny"] * a["Nz"], size=n_pix, replace=False)
    threads, host = host_cpu()
    kw = dict(pos_m=a["pos_m"], f0p=a["f0p"], f0z=a["f0z"], sanr=a["sanr"], sani=a["sani"],
              amp_r=a["amp_r"], amp_i=a["amp_i"], cx=a["cx"][sel], cy=a["cy"][sel],
              cz=a["cz"][sel], k_val=a["k_val"], remove_forward=False, Ny=n_pix, Nz=1,
              apply_pol=True, pol_deg=a["pol_deg"], pol_e=a["pol_e"], apply_decay=False)
    O.ref_kin_sum(**{**kw, "pos_m": a["pos_m"][:2000], "f0p": a["f0p"][:2000],
                     "f0z": a["f0z"][:2000], "sanr": a["sanr"][:2000], "sani": a["sani"][:2000],
                     "amp_r": a["amp_r"][:2000], "amp_i": a["amp_i"][:2000]},
                  threads=threads)  # warm-up
    t = time.perf_counter()
    O.ref_kin_sum(**kw, threads=threads)
    dt = time.perf_counter() - t
    pairs = a["pos_m"].shape[0] * n_pix
    return {"value": pairs / dt, "unit": "atom*pixel pairs/s", "cores": threads,
            "kind": "reference", "host": host,
            "sample": f"compute_scattering_cffi (oracle/_ref, gcc -O3 -march=x86-64-v3) on all "
                      f"{a['pos_m'].shape[0]} atoms x {n_pix} random pixels, "
                      f"{threads} threads over >=1000-atom slices, {dt:.1f} s"}


def accuracy(w, field, n_pix, seed=1):
    """Relative error of the GPU field against the FP64 oracle on random pixels."""
    from mosaic_amd import workloads as W
    from oracle import kinematic as O
    a = W.oracle_inputs(w)
    rng = np.random.default_rng(seed)
    sel = rng.choice(a["Ny"] * a["Nz"], size=n_pix, replace=False)
    R = O.kin_sum_c(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                    a["amp_i"], a["cx"][sel], a["cy"][sel], a["cz"][sel], a["k_val"], False,
                    n_pix, 1, True, a["pol_deg"], a["pol_e"], False, precision="f64",
                    threads=host_cpu()[0])
    A = field.ravel()[sel]
    return float(np.linalg.norm(A - R) / np.linalg.norm(R))


def bench_exact(reps=3, n_err=24):
    """The exact per-pair forms on the surrogate workload's crystal (FCC Au,
    37^3 cells, 202,612 atoms) and a 128^2 detector at its (111) Bragg angle:
    the staged sum forced onto the exact kernel (the reference's general
    kernel, interaction_kernal, in its exact mode) and mosaic_scatter
    (compute_scattering_cffi on the device: lab positions in FP64 metres,
    per-atom table rows), each with its error against the FP64 oracle on
    random pixels.  Inputs resident in HBM; HIP events around each call."""
    import torch
    from mosaic_amd import kinematic as K
    from mosaic_amd import workloads as W
    from oracle import kinematic as O
    w = W.fcc_fault_scan111(cells=(37, 37, 37), npix=128, n_angles=1, fault=False)
    b, smp, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    a = W.oracle_inputs(w)
    n = a["pos_m"].shape[0]
    Ny, Nz = a["Ny"], a["Nz"]
    dev = torch.device("cuda", 0)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    # staged sum, exact kernel
    b._kin_path = "exact"
    rot = np.asarray(st.rotation, dtype=np.float32)[None]
    tr = np.asarray(st.translation, dtype=np.float32)[None]
    f = b._kin_field(smp, d.pixel_coordinates, d.shape, st, (rot, tr), False, True, False)
    torch.cuda.synchronize()
    field = f[0].cpu().numpy().ravel()
    # time the sum alone: the staged sample of the call above
    pos = torch.as_tensor(np.concatenate([smp.load_chunk_positions(c) for c in
                                          range(1, smp.chunk_total + 1)]), device=dev)
    codes = torch.zeros(n, dtype=torch.uint8, device=dev)
    import mosaic_amd.tables as T
    spc = K.species_table(T.species_constants(["Au"], b.energy))
    opts = K.kin_opts(a["k_val"], False, True, b._polarization_state(), False, path="exact")
    staged = K.stage(pos, codes, None, np.array([n]), a["k_val"], 1, Ny, Nz)
    coords = torch.as_tensor(np.ascontiguousarray(np.asarray(d.pixel_coordinates) / 1e10),
                             device=dev)
    out = torch.zeros((1, Nz, Ny), dtype=torch.complex64, device=dev)
    K.kin_sum(staged, spc, opts, rot, tr, coords, out)
    ev[0].record()
    for _ in range(reps):
        K.kin_sum(staged, spc, opts, rot, tr, coords, out)
    ev[1].record()
    # mosaic_scatter on the oracle's own per-atom arrays
    t = {k: torch.as_tensor(np.ascontiguousarray(a[k]), device=dev)
         for k in ("pos_m", "f0p", "f0z", "sanr", "sani", "amp_r", "amp_i", "cx", "cy", "cz")}
    t["pos_m"] = t["pos_m"].double()
    for k in ("f0p", "f0z", "sanr", "sani", "amp_r", "amp_i"):
        t[k] = t[k].float()
    for k in ("cx", "cy", "cz"):
        t[k] = t[k].double()
    o_r = torch.zeros(Ny * Nz, dtype=torch.float32, device=dev)
    o_i = torch.zeros_like(o_r)
    args = (t["pos_m"], t["f0p"], t["f0z"], t["sanr"], t["sani"], t["amp_r"], t["amp_i"],
            t["cx"], t["cy"], t["cz"], a["k_val"], False, Ny, Nz, True, a["pol_deg"],
            a["pol_e"], False)
    K.scatter(*args, o_r, o_i)
    torch.cuda.synchronize()
    sc_field = o_r.cpu().numpy().astype(np.float64) + 1j * o_i.cpu().numpy()
    ev[2].record()
    for _ in range(reps):
        K.scatter(*args, o_r, o_i)
    ev[3].record()
    ev[3].synchronize()
    t_ex = ev[0].elapsed_time(ev[1]) / reps
    t_sc = ev[2].elapsed_time(ev[3]) / reps
    rng = np.random.default_rng(5)
    sel = rng.choice(Ny * Nz, size=n_err, replace=False)
    R = O.kin_sum_c(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                    a["amp_i"], a["cx"][sel], a["cy"][sel], a["cz"][sel], a["k_val"], False,
                    n_err, 1, True, a["pol_deg"], a["pol_e"], False, precision="f64",
                    threads=host_cpu()[0])
    pairs = float(n) * Ny * Nz
    return {"workload": "fcc_au_37c_128px (202,612 atoms x 128^2 pixels)",
            "exact_kernel_ms": t_ex, "exact_kernel_pairs_per_s": pairs / (t_ex * 1e-3),
            "exact_rel_err_vs_fp64_oracle": float(np.linalg.norm(field[sel] - R) /
                                                  np.linalg.norm(R)),
            "scatter_ms": t_sc, "scatter_pairs_per_s": pairs / (t_sc * 1e-3),
            "scatter_rel_err_vs_fp64_oracle": float(np.linalg.norm(sc_field[sel] - R) /
                                                    np.linalg.norm(R)),
            "rel_err_sample": f"{n_err} random pixels, all atoms"}


PROP_N = 2048           # exit-wave grid
PROP_PITCH = 170e-9     # m
HBM_PEAK = 8.0e12       # B/s


def bench_propagation(steps, warmup, wavelength):
    """CRL (f = 0.5 m) then 2 m of free space on a 2048^2 complex64 exit wave at
    170 nm pitch: two alias-free 1 m steps on 8192^2 padded grids.

    Algorithmic HBM bytes per step of the pruned transform (the default):
    16 Ny Nz + 32 Ny2 Nz -- the field read and written once, and the
    Ny2 x Nz intermediate (rows transformed along y) written, read and
    rewritten in place, and read again.  The padded 2-D FFT round trip is
    timed beside it; its own minimum would be 64 Ny2 Nz2 (two 2-pass FFTs
    over the padded grid)."""
    import torch
    from mosaic_amd import propagation as MP
    from mosaic_amd.optics import optics
    dev = torch.device("cuda", 0)
    n, d = PROP_N, PROP_PITCH
    x = (np.arange(n) - (n - 1) / 2.0) * d
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    E0 = (np.exp(-r2 / (0.3 * n * d) ** 2) * np.exp(1j * 2e9 * r2)).astype(np.complex64)
    E = torch.from_numpy(E0).to(dev)
    ops = optics()
    ops.add_CRL_box(1, 500.0, 1.0)
    ops.add_free_space(2000.0)
    plan = MP.plan_steps(n, n, d, d, wavelength, 2.0)

    def timed(method):
        def prop_fn(F, dym, dzm, z, lens=None):
            return MP.propagate(F, dym, dzm, wavelength, z, lens=lens, method=method)
        prop_fn.accepts_lens = True
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(max(1, warmup)):
            ops.apply_stack_device(E, d, d, wavelength, prop_fn)
        ts = []
        for _ in range(steps):
            ev0.record()
            out = ops.apply_stack_device(E, d, d, wavelength, prop_fn)
            ev1.record()
            ev1.synchronize()
            ts.append(ev0.elapsed_time(ev1))
        return float(np.mean(ts)), out

    ms, out = timed("auto")
    ms2d, out2d = timed("fft2d")
    alg = sum(16.0 * n * n + 32.0 * a * n for _, a, b in plan)
    alg2d = sum(64.0 * a * b for _, a, b in plan)
    return {"workload": "CRL f=0.5 m + 2 m free space, 2048^2 complex64 at 170 nm",
            "method": "pruned three-pass transform",
            "substeps": [[float(z), a, b] for z, a, b in plan], "ms": ms,
            "algorithmic_bytes": alg, "achieved_GBps": alg / (ms * 1e-3) / 1e9,
            "peak_GBps": HBM_PEAK / 1e9, "frac": alg / (ms * 1e-3) / HBM_PEAK,
            "traffic_per_step_bytes": sum((2.0 * f + w_) * 1024.0
                                          for f, w_ in PROP_COUNTERS.values()),
            "traffic_source": "profiles/propagation_crl2m_2048_on_8192.md",
            "padded_fft2d_ms": ms2d,
            "padded_fft2d_frac": alg2d / (ms2d * 1e-3) / HBM_PEAK,
            "rel_diff_vs_padded": float(torch.linalg.norm(out - out2d) / torch.linalg.norm(out2d)),
            "out_energy": float(torch.sum(torch.abs(out) ** 2).item())}


def bench_scan(steps, warmup):
    """Batched rocking scan: 1,000,000-atom FCC Au with a (111) stacking fault,
    64 rotations in +-0.05 deg, 256^2 pixels, all poses in one launch."""
    import torch
    from mosaic_amd import kinematic as K
    from mosaic_amd import tables
    from mosaic_amd import workloads as W
    w = W.fcc_fault_scan111(cells=(50, 50, 100), npix=256, n_angles=64)
    b, d, st = w["beam"], w["detector"], w["stage"]
    dev = torch.device("cuda", 0)
    pos = torch.from_numpy(w["sample"].load_chunk_positions(1)).to(dev)
    n = pos.shape[0]
    codes = torch.zeros(n, dtype=torch.uint8, device=dev)
    rots, trs = [], []
    for a in w["angles_deg"]:
        st.set_single_motor_value_absolute("phi", float(a))
        rots.append(np.asarray(st.rotation, np.float32))
        trs.append(np.asarray(st.translation, np.float32))
    st.set_single_motor_value_absolute("phi", 0.0)
    rot, tr = np.stack(rots), np.stack(trs)
    F = rot.shape[0]
    Ny, Nz = d.shape
    k = 2 * np.pi / b.wavelength
    spc = K.species_table(tables.species_constants(["Au"], b.energy))
    opts = K.kin_opts(k, False, True, b._polarization_state(), False)
    coords = torch.from_numpy(np.ascontiguousarray(d.pixel_coordinates / 1e10)).to(dev)
    out = torch.zeros((F, Nz, Ny), dtype=torch.complex64, device=dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ws = [None]

    def run():
        out.zero_()
        sg = K.stage(pos, codes, None, np.array([n]), k, F, Ny, Nz, workspace=ws[0])
        ws[0] = sg.workspace
        K.kin_sum(sg, spc, opts, rot, tr, coords, out)
        return sg

    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(steps):
        ev0.record()
        sg = run()
        ev1.record()
        ev1.synchronize()
        ts.append(ev0.elapsed_time(ev1))
    ms = float(np.mean(ts))
    pairs = float(n) * Ny * Nz * F
    # error: 3 frames x 32 random pixels against the FP64 oracle over all atoms
    from oracle import kinematic as O
    field = out.cpu().numpy().reshape(F, -1)
    rng = np.random.default_rng(21)
    num = den = 0.0
    for f in (0, F // 2, F - 1):
        a = W.oracle_inputs(w, frame_rot=rot[f], frame_trans=tr[f])
        sel = rng.choice(Ny * Nz, size=32, replace=False)
        R = O.kin_sum_c(a["pos_m"], a["f0p"], a["f0z"], a["sanr"], a["sani"], a["amp_r"],
                        a["amp_i"], a["cx"][sel], a["cy"][sel], a["cz"][sel], a["k_val"], False,
                        32, 1, True, a["pol_deg"], a["pol_e"], False, precision="f64",
                        threads=host_cpu()[0])
        num += float(np.sum(np.abs(field[f][sel] - R) ** 2))
        den += float(np.sum(np.abs(R) ** 2))
    return {"workload": w["name"], "atoms": n, "frames": F, "pixels": [Ny, Nz],
            "ms_per_scan": ms, "frames_per_s": F / (ms * 1e-3),
            "pairs_per_s": pairs / (ms * 1e-3), "kernel_path": sg.path_used,
            "rel_err_vs_fp64_oracle": float(np.sqrt(num / den)),
            "rel_err_sample": "frames 0, 31, 63 x 32 random pixels, all atoms"}


def bench_end_to_end(w, steps):
    """The drop-in as a MOSAIC caller uses it: beam.atomic_direct_interaction
    on the host-resident chunked sample (64 chunks of the 5M-atom workload):
    chunks loaded and copied from pinned host memory, incident amplitudes,
    k-d staging, pixel prologue, sum, the field copied back and written to
    the detector.  Host clock around each call, device synchronized."""
    import torch
    b, s, d, st = w["beam"], w["sample"], w["detector"], w["stage"]
    kw = dict(apply_polarization=True)
    b.atomic_direct_interaction(s, d, st, sc_kwargs=kw)      # warm-up (pinned buffers)
    torch.cuda.synchronize()
    ts = []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        b.atomic_direct_interaction(s, d, st, sc_kwargs=kw)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ms = float(np.median(ts)) * 1e3
    # the same sample in page-locked memory (sample.pin_memory()): the chunks
    # go to the device by DMA, without the host copy into the staging buffer
    s.pin_memory()
    b.atomic_direct_interaction(s, d, st, sc_kwargs=kw)
    torch.cuda.synchronize()
    tp = []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        b.atomic_direct_interaction(s, d, st, sc_kwargs=kw)
        torch.cuda.synchronize()
        tp.append(time.perf_counter() - t0)
    ms_p = float(np.median(tp)) * 1e3
    n = s.atom_count
    Ny, Nz = d.shape
    return {"call": "beam.atomic_direct_interaction(sample, detector, stage, "
                    "sc_kwargs={'apply_polarization': True})",
            "chunks": s.chunk_total, "ms": ms,
            "pairs_per_s": float(n) * Ny * Nz / (ms * 1e-3),
            "pinned_sample_ms": ms_p,
            "pinned_sample_pairs_per_s": float(n) * Ny * Nz / (ms_p * 1e-3),
            "h2d_bytes": 13 * n, "d2h_bytes": 8 * Ny * Nz,
            "kernel_path": b._last_path}


def roofline(path_used, pairs, ksum_ms, workload, kernel_ms=None):
    """Roofline object of the dominant (sum) kernel.

    ``achieved`` is the work the kernel's own algorithm does per launch over
    its live launch time: for the VALU kernels the FP32 flops it executes
    (add, mul, 2 per fma, one per transcendental; 64 lanes per counted
    wave-instruction, from the profile's SQ_INSTS_VALU_*_F32 counters of the
    same workload -- the counts are fixed by the workload), against the
    157.3 TFLOP/s FP32 peak; for the matrix-core kernel the FP16 matrix
    flops it issues, against 2.5 PFLOP/s (its FP32 phase products, 2 x 32 x
    32 x 2 flop each, are reported apart: ``mfma_f32_flop_per_launch``).
    ``pair_equivalent_tflops`` is a
    throughput figure, not a fraction of a roofline: 20 flop per pair (the
    per-pair algorithm) times the pairs the kernel's result stands for."""
    name = kernel_name(path_used)
    c = COUNTERS.get(name) if workload == WORKLOAD_OF_COUNTERS else None
    if name == "k_gm_moments" and kernel_ms is not None:
        # the matrix-core moment sum is two kernels: the moments (dominant)
        # and their evaluation, timed apart by the library's HIP events
        t = kernel_ms[0] * 1e-3
        r = {"kernel": name, "kernel_path": path_used, "kernel_ms": kernel_ms[0],
             "eval_kernel": "k_gm_eval", "eval_ms": kernel_ms[1], "sum_ms": ksum_ms,
             "eval_pixel_fallbacks": int(round(kernel_ms[2])) if len(kernel_ms) > 2 else None,
             "sum_pairs_per_s": pairs / (ksum_ms * 1e-3), "unit": "TFLOP/s",
             "pair_equivalent_tflops": pairs * FLOP_PER_PAIR / (ksum_ms * 1e-3) / 1e12,
             "traffic": None}
        if c is not None:
            r["traffic"] = (2.0 * c["fetch_kib"] + c["write_kib"]) * 1024.0
            r["counters_source"] = c["source"]
            flop = c["mfma"] * MFMA_FLOP_PER_INST
            v32 = 64.0 * (c["add_f32"] + c["mul_f32"] + 2.0 * c["fma_f32"] + c["trans_f32"])
            r.update({"bound": "mfma", "achieved": flop / t / 1e12,
                      "peak": PEAK_FP16_MFMA / 1e12, "frac": flop / t / PEAK_FP16_MFMA,
                      "mfma_flop_per_launch": flop,
                      "mfma_f32_flop_per_launch": c.get("mfma_f32", 0.0) * MFMA_F32_FLOP_PER_INST,
                      "fp32_valu_tflops": v32 / t / 1e12, "fp32_valu_frac": v32 / t / PEAK_FP32,
                      "valu_issue_frac": 64.0 * c["valu"] / t / PEAK_VALU_LANE_OPS,
                      "trans_frac": 64.0 * c["trans_f32"] / t / PEAK_TRANS_LANE_OPS})
        else:
            r.update({"bound": "mfma", "peak": PEAK_FP16_MFMA / 1e12, "achieved": None,
                      "frac": None, "note": "no counters of this kernel on this workload"})
        return r
    t = ksum_ms * 1e-3
    r = {"kernel": name, "kernel_path": path_used, "kernel_ms": ksum_ms,
         "kernel_pairs_per_s": pairs / t, "unit": "TFLOP/s",
         "pair_equivalent_tflops": pairs * FLOP_PER_PAIR / t / 1e12,
         "traffic": None}
    if c is not None:
        r["traffic"] = (2.0 * c["fetch_kib"] + c["write_kib"]) * 1024.0
        r["counters_source"] = c["source"]
    if name == "k_sum_mma":
        flop = pairs * MFMA_FLOP_PER_PAIR
        r.update({"bound": "mfma", "achieved": flop / t / 1e12, "peak": PEAK_FP16_MFMA / 1e12,
                  "frac": flop / t / PEAK_FP16_MFMA})
        return r
    r.update({"bound": "valu", "peak": PEAK_FP32 / 1e12})
    if c is not None and c.get("fma_f32") is not None:
        flop = 64.0 * (c["add_f32"] + c["mul_f32"] + 2.0 * c["fma_f32"] + c["trans_f32"])
        r.update({"achieved": flop / t / 1e12, "frac": flop / t / PEAK_FP32,
                  "fp32_flop_per_launch": flop, "fp32_flop_per_pair": flop / pairs,
                  "trans_lane_ops_per_s": 64.0 * c["trans_f32"] / t,
                  "trans_frac": 64.0 * c["trans_f32"] / t / PEAK_TRANS_LANE_OPS,
                  "valu_lane_ops_per_pair": 64.0 * c["valu"] / pairs,
                  "valu_issue_frac": 64.0 * c["valu"] / t / PEAK_VALU_LANE_OPS})
    else:
        r.update({"achieved": None, "frac": None,
                  "note": "no counters of this kernel on this workload"})
    return r


def kernel_name(path_used):
    """Sum kernel behind a mosaic_kin_sum path code."""
    base = path_used % 100
    return {32: "k_sum_mma", 24: "k_sum_rec", 28: "k_sum_rec", 45: "k_sum_poly",
            46: "k_sum_poly", 76: "k_gm_moments", 50: "k_sum_exact", 51: "k_sum_exact",
            52: "k_sum_exact"}.get(base, "k_sum")


def bench_adjoint(steps):
    """Displacement gradient: 1,000,188-atom FCC Au (63^3 cells) with a Gaussian
    strain bump, 16 poses x 256^2 pixels, L = sum ((|A|^2 - I_t) / max I_t)^2."""
    import torch
    from mosaic_amd import workloads as W
    from mosaic_amd.fitting import KinematicModel
    w = W.fcc_strain_fit(cells=(63, 63, 63), npix=256, n_angles=16)
    b, d, st = w["beam"], w["detector"], w["stage"]
    dev = torch.device("cuda", 0)
    pos = w["sample"].load_chunk_positions(1)
    rots, trs = [], []
    for a in w["angles_deg"]:
        st.set_single_motor_value_absolute("phi", float(a))
        rots.append(np.asarray(st.rotation, np.float32))
        trs.append(np.asarray(st.translation, np.float32))
    st.set_single_motor_value_absolute("phi", 0.0)
    m = KinematicModel(pos, np.zeros(pos.shape[0], np.uint8), ["Au"], b, d, np.stack(rots),
                       np.stack(trs))
    with torch.no_grad():
        It = (m.field(torch.from_numpy(w["u_target"]).to(dev)).abs() ** 2)
    scale = float(It.max())
    u = torch.from_numpy(w["u"]).to(dev).requires_grad_(True)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    tf, tb = [], []
    for i in range(steps + 1):
        u.grad = None
        ev[0].record()
        A = m.field(u)
        L = torch.sum(((A.abs() ** 2 - It) / scale) ** 2)
        ev[1].record()
        L.backward()
        ev[2].record()
        ev[2].synchronize()
        if i:
            tf.append(ev[0].elapsed_time(ev[1]))
            tb.append(ev[1].elapsed_time(ev[2]))
    pairs = float(pos.shape[0]) * 16 * 256 * 256
    # error: dL/du of 20 random atoms against the oracle's reference gradient
    # (exact per-pair phase gradient, long-double atom terms), given the
    # GPU field's upstream gradient
    from mosaic_amd import tables
    from oracle import adjoint as AO
    grad = u.grad.cpu().numpy()
    A_np = A.detach().cpu().numpy().reshape(16, -1)
    It_np = It.cpu().numpy().reshape(16, -1)
    g = 4 * (np.abs(A_np) ** 2 - It_np) * A_np / scale ** 2
    c = tables.species_constants(["Au"], b.energy)
    k = 2 * np.pi / b.wavelength
    pc = d.pixel_coordinates / 1e10
    rot_a, tr_a = np.stack(rots), np.stack(trs)
    p_all = (pos + w["u"]).astype(np.float32)
    atoms = np.random.default_rng(31).choice(pos.shape[0], size=20, replace=False)
    ref = []
    for j in atoms:
        cj = AO.atom_terms(p_all[j], c["wk"][0], float(c["f0z"][0]), float(c["fp"][0]),
                           float(c["fpp"][0]), np.complex64(1.0), rot_a, tr_a, pc, k,
                           b._polarization_state(), host_cpu()[0])
        ref.append(AO.reference_gradient(cj, AO.phase_gradient(p_all[j], rot_a, tr_a, pc, k),
                                         g))
    ref = np.array(ref)
    e = float(np.linalg.norm(grad[atoms] - ref) / np.linalg.norm(ref))
    return {"workload": w["name"], "atoms": int(pos.shape[0]), "poses": 16,
            "forward_ms": float(np.mean(tf)), "backward_ms": float(np.mean(tb)),
            "backward_pairs_per_s": pairs / (np.mean(tb) * 1e-3),
            "grad_norm": float(torch.linalg.norm(u.grad).item()),
            "rel_err_vs_reference_gradient": e,
            "rel_err_sample": "20 random atoms, long-double reference gradient"}


# multiply-accumulates of one member on one image: the MLP 11-256-512-4096
# and the four 2x2 stride-2 transposed convolutions (input pixels x Cin x 4 Cout)
SG_MACS = (11 * 256 + 256 * 512 + 512 * 4096 + 64 * 64 * 256 + 256 * 64 * 128 +
           1024 * 32 * 64 + 4096 * 16 * 4)


def bench_surrogate_inference(reps, members=5, frames=64):
    """The surrogate ensemble's fused inference (mosaic_sg_infer, two
    launches) on one 64-angle scan, against its layer-by-layer form.
    Random-init weights of the trained architecture: the work does not depend
    on their values."""
    import torch
    from mosaic_amd import surrogate as SG
    model = SG.SurrogateEnsemble(members=members, seed=0)
    rng = np.random.default_rng(0)
    X = torch.as_tensor(rng.uniform(-1, 1, (frames, SG.N_FEAT)).astype(np.float32),
                        device=model.device)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for _ in range(3):
        model.infer(X)
        model.forward(X)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        model.infer(X)
    ev[1].record()
    for _ in range(reps):
        model.forward(X)
    ev[2].record()
    ev[2].synchronize()
    t_f = ev[0].elapsed_time(ev[1]) / reps
    t_l = ev[1].elapsed_time(ev[2]) / reps
    flop = 2.0 * SG_MACS * members * frames
    return {"members": members, "frames": frames, "fused_us": t_f * 1e3,
            "layered_us": t_l * 1e3, "flop": flop,
            "achieved_tflops": flop / (t_f * 1e-3) / 1e12, "peak_tflops_bf16": 2500.0,
            "frac": flop / (t_f * 1e-3) / 2.5e15,
            "note": "algorithmic BF16 MFMA flops (15.5 MFLOP per member-image) over the "
                    "HIP-event time of both launches; the elementwise SiLU work (v_exp, "
                    "v_rcp per activation) and launch latency bound it, not the matrix cores"}


def bench_surrogate(n_images, epochs, members=5, seed=0, n_scans=8, judge_pixels=64,
                    threshold=1e-2):
    """Learned rocking scan: ensemble trained on GPU-generated images of a
    200k-atom FCC Au crystal (37^3 cells) with a random dislocation, strain
    and stacking fault; 64-angle scans anchored on plain frames and judged
    by exact probes of every other frame.  Beside it: the network alone, the
    anchors alone (linear interpolation between the same two plain frames;
    quadratic interpolation through three), i.e. the method without the
    network, the fused inference's time, and the floor that the anchors and
    the probe set on any judged scan."""
    import torch
    from mosaic_amd import surrogate as SG
    t0 = time.perf_counter()
    gen = SG.ImageGenerator(cells=(37, 37, 37), npix=128)
    P, I = SG.make_training_set(gen, n_images, seed=seed)
    Pv, Iv = SG.make_training_set(gen, max(256, n_images // 20), seed=seed + 1,
                                  report_every=0)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    model = SG.SurrogateEnsemble(members=members, seed=seed)
    t1 = time.perf_counter()
    hist = model.train(P, I, epochs=epochs, batch=128, lr=2e-3, seed=seed,
                       log=lambda m: print(m, file=sys.stderr, flush=True))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    Ipv, _ = model.predict(Pv)
    errv = SG.rel_l2(Ipv, Iv).cpu().numpy()
    angles = np.linspace(-0.05, 0.05, 64)
    rng = np.random.default_rng(seed + 7)
    scans = []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(10)]
    scan = SG.SurrogateScan(model, gen, threshold, judge_pixels=judge_pixels, seed=seed)
    inner = np.arange(1, 63)
    for sidx in range(n_scans):
        p = SG.sample_params(1, rng)[0]
        gen.intensities(p, angles[:2])          # warm
        scan.run(p, angles)
        ev[0].record()
        Ip = gen.intensities(p, angles)
        ev[1].record()
        ev[2].record()
        Is, est, bad = scan.run(p, angles)
        ev[3].record()
        # the parts: fused inference, anchors, probes
        P64 = np.column_stack([np.repeat(p[None, :8], 64, 0), angles])
        ev[4].record()
        model.member_intensities(P64)
        ev[5].record()
        gen.intensities(p, angles[[0, 63]])
        ev[6].record()
        pix = torch.randint(0, Ip.shape[1], (judge_pixels,), device=Ip.device)
        ev[7].record()
        gen.probe(p, angles[inner], pix)
        ev[8].record()
        ev[8].synchronize()
        t_plain = ev[0].elapsed_time(ev[1])
        t_sg = ev[2].elapsed_time(ev[3])
        err = SG.rel_l2(Is, Ip).cpu().numpy()
        est_c = est.cpu().numpy()
        acc_mask = np.ones(64, bool)
        acc_mask[list(bad)] = False
        acc_mask[[0, 63]] = False
        # the judge's own quality on this scan: the unjudged frames' true error
        # against its estimate (before any fallback)
        W = torch.as_tensor(SG.anchor_weights(angles, angles[[0, 63]]), device=Ip.device)
        Im = model.member_intensities(P64)
        pre = (Im + torch.matmul(W, Ip[[0, 63]][None] - Im[:, [0, 63]])).mean(0).clamp_min(0)
        pre_err = SG.rel_l2(pre, Ip).cpu().numpy()[inner]
        raw = SG.rel_l2(model.predict(P64)[0], Ip).cpu().numpy()
        # the anchors without the network
        W = torch.as_tensor(SG.anchor_weights(angles, angles[[0, 63]]), device=Ip.device)
        lin = SG.rel_l2(torch.matmul(W, Ip[[0, 63]]), Ip).cpu().numpy()
        x = angles[[0, 32, 63]]
        Lq = np.stack([np.prod([(angles - x[m]) / (x[j] - x[m]) for m in range(3) if m != j],
                               axis=0) for j in range(3)], axis=1).astype(np.float32)
        quad = SG.rel_l2(torch.matmul(torch.as_tensor(Lq, device=Ip.device), Ip[[0, 32, 63]]),
                         Ip).cpu().numpy()
        sp = scan.last_spread.cpu().numpy()
        scans.append({"plain_ms": t_plain, "surrogate_ms": t_sg,
                      "speedup": t_plain / t_sg, "fallback_frames": int(len(bad)),
                      "spread_fallbacks": int(np.sum(sp > scan.spread_limit)),
                      "spread_max": float(sp.max()),
                      "inference_ms": ev[4].elapsed_time(ev[5]),
                      "anchors_ms": ev[5].elapsed_time(ev[6]),
                      "probe_ms": ev[7].elapsed_time(ev[8]),
                      "max_rel_err": float(err.max()), "median_rel_err": float(np.median(err)),
                      "accepted_max_rel_err": float(err[acc_mask].max()) if acc_mask.any()
                      else 0.0,
                      "judge_est_vs_true_corr": float(np.corrcoef(est_c[inner], pre_err)[0, 1]),
                      "judge_est_over_true_median": float(np.median(est_c[inner] / pre_err)),
                      "network_alone_median_rel_err": float(np.median(raw)),
                      "network_alone_max_rel_err": float(raw.max()),
                      "anchors_linear_median_rel_err": float(np.median(lin)),
                      "anchors_linear_max_rel_err": float(lin.max()),
                      "anchors_quadratic3_max_rel_err": float(quad.max())})
        print(f"[surrogate] scan {sidx}: {json.dumps(scans[-1])}", file=sys.stderr, flush=True)

    def agg(key, f=np.median):
        return float(f([x[key] for x in scans]))
    return {"workload": "fcc_au_37c_128px defect scan, 64 angles in +-0.05 deg",
            "images": n_images, "members": members, "epochs": epochs,
            "params_per_member": model.n_params(), "target": model.target,
            "gen_s": t_gen, "train_s": t_train,
            "final_loss": [float(x) for x in hist[-1]], "judge_threshold": threshold,
            "judge_pixels": judge_pixels, "spread_limit": scan.spread_limit,
            "spread_fallbacks_total": int(sum(x["spread_fallbacks"] for x in scans)),
            "val_median_rel_err": float(np.median(errv)),
            "val_frac_below_1e-2": float(np.mean(errv <= 1e-2)),
            "plain_frames_per_s": 64 / (np.mean([x["plain_ms"] for x in scans]) * 1e-3),
            "surrogate_frames_per_s": 64 / (np.mean([x["surrogate_ms"] for x in scans]) * 1e-3),
            "speedup": float(np.mean([x["plain_ms"] for x in scans]) /
                             np.mean([x["surrogate_ms"] for x in scans])),
            "scan_max_rel_err": agg("max_rel_err", np.max),
            "accepted_max_rel_err": agg("accepted_max_rel_err", np.max),
            "inference_ms": agg("inference_ms"), "anchors_ms": agg("anchors_ms"),
            "probe_ms": agg("probe_ms"),
            "floor_speedup": float(np.mean([x["plain_ms"] for x in scans]) /
                                   (agg("anchors_ms") + agg("probe_ms") + agg("inference_ms"))),
            "judge_est_vs_true_corr": agg("judge_est_vs_true_corr"),
            "scan_median_rel_err": agg("median_rel_err"),
            "fallback_frames_total": int(sum(x["fallback_frames"] for x in scans)),
            "network_alone_median_rel_err": agg("network_alone_median_rel_err"),
            "anchors_linear_max_rel_err": agg("anchors_linear_max_rel_err", np.max),
            "anchors_quadratic3_max_rel_err": agg("anchors_quadratic3_max_rel_err", np.max),
            "scans": scans}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the kinematic sum is "
              "not sharded across GPUs in this build", file=sys.stderr)
        sys.exit(2)
    import torch
    from mosaic_amd import kinematic as K
    from mosaic_amd import tables

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.sg_infer_only:
        print(json.dumps({"surrogate_inference": bench_surrogate_inference(50)}), flush=True)
        return
    if args.surrogate:
        print(json.dumps({"surrogate": bench_surrogate(args.sg_images, args.sg_epochs,
                                                       n_scans=args.sg_scans)}),
              flush=True)
        return
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)      # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_copy": 1 << 30,
                                               "bytes_written_per_copy": 1 << 30,
                                               "copies": 3}}), flush=True)
        return
    w = build_workload(args.workload)
    b, d, st = w["beam"], w["detector"], w["stage"]
    pos, codes, labels = gather(w)
    n_atoms = pos.shape[0]
    Ny, Nz = d.shape
    npix = Ny * Nz
    pairs = float(n_atoms) * npix
    k_val = 2.0 * np.pi / b.wavelength
    consts = tables.species_constants(labels, b.energy)
    spc = K.species_table(consts)
    opts = K.kin_opts(k_val, False, True, b._polarization_state(), False, path=args.path)
    rot = np.asarray(st.rotation, dtype=np.float32)[None]
    tr = np.asarray(st.translation, dtype=np.float32)[None]
    counts = np.bincount(codes, minlength=len(labels)).astype(np.int64)

    d_pos = torch.from_numpy(pos).to(dev)
    d_cod = torch.from_numpy(codes).to(dev)
    coords = torch.from_numpy(np.ascontiguousarray(d.pixel_coordinates / 1e10)).to(dev)
    E0 = torch.from_numpy(b._beam_E0_map).to(dev)
    amp = K.ein_bilinear(d_pos, rot[0], tr[0], E0, b._beam_du, b._beam_dv, b._beam_uc,
                         b._beam_vc)
    out = torch.zeros((1, Nz, Ny), dtype=torch.complex128, device=dev)
    ws = [None]
    # HIP events around each timed step's sum, read after the timed loop (no
    # host wait inside it beyond the staging call's own)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(max(1, args.steps))]

    path_used = [None]

    def step(i=None):
        out.zero_()
        staged = K.stage(d_pos, d_cod, amp, counts, k_val, 1, Ny, Nz, workspace=ws[0])
        ws[0] = staged.workspace
        K.kin_pixels(staged, spc, opts, coords)
        if i is not None:
            evs[i][0].record()
        K.kin_sum(staged, spc, opts, rot, tr, coords, out)
        if i is not None:
            evs[i][1].record()
        path_used[0] = staged.path_used

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    K.sum_timing(True)          # a fresh record of the timed steps' kernel launches
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sum_ms = [a.elapsed_time(b) for a, b in evs[:args.steps]]
    kern_ms = K.sum_timing(False, fallbacks=True)   # means over the timed steps
    for i, v in enumerate(sum_ms):
        print(f"[bench] step {i + 1}/{args.steps} k_sum {v:.2f} ms", file=sys.stderr, flush=True)
    ms_step = (t1 - t0) * 1e3 / args.steps
    ksum_ms = float(np.mean(sum_ms))
    field = out[0].cpu().numpy()
    res = {
        "metric": "atom*pixel pairs/s (kinematic sum)",
        "value": pairs / (ms_step * 1e-3),
        "unit": "atom*pixel pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "atoms": int(n_atoms), "pixels": [Ny, Nz],
                   "pairs_per_step": pairs, "species": labels, "energy_eV": b.energy},
        "roofline": roofline(path_used[0], pairs, ksum_ms, args.workload,
                             tuple(kern_ms)),
    }
    if not args.no_end_to_end:
        res["end_to_end"] = bench_end_to_end(w, args.steps)
    if not args.no_propagation:
        res["propagation"] = bench_propagation(args.steps, args.warmup, b.wavelength)
    if not args.no_scan:
        res["scan"] = bench_scan(max(1, args.steps - 1), 1)
    if not args.no_adjoint:
        res["adjoint"] = bench_adjoint(1)
    if not args.no_surrogate_inference:
        res["surrogate_inference"] = bench_surrogate_inference(max(3, args.steps))
    if not args.no_exact:
        res["exact"] = bench_exact()
    if args.err_pixels > 0:
        res["rel_err_vs_fp64_oracle"] = accuracy(w, field, args.err_pixels)
        res["rel_err_pixels"] = args.err_pixels
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(w, args.cpu_pixels)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
