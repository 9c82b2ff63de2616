#!/usr/bin/env python3
"""Time to the 50 largest singular triplets of a 2,000,000 x 1,000,000 sparse
matrix (~32M nonzeros) on one MI355X, with the achieved HBM bandwidth of the
dominant kernel.

Workload (one "step" = one full restarted solve):
  A: 2,000,000 x 1,000,000 CSR, int32 indices, each row Poisson(16) distinct
     uniform columns, value N(0,1) * (1 + j)^-0.6 for column j (seeded).
  DLANSVD_IRL: k = 50 largest triplets, Lanczos dimension 120, p = 70 shifts,
  tol = 1e-10, delta = sqrt(eps), eta = eps^(3/4), ICGS, ELR, seeded N(0,1)
  start vector; singular vectors computed (left on the device).
The matrix (device CSR operator, SpMV layout from the autotuner) and the start
vector are resident in HBM before timing starts.

Also reported: live HIP-event timing of the SpMV (the roofline kernel) inside
the timed solves, every kernel class timed in one more solve after them, the
SpMV and reorthogonalization bandwidth together, the max relative residual of
the 50 triplets, the FP32-basis form with FP64 polish on the same matrix, and
the reference PROPACK (OpenMP build, oracle/_ref) timed on the host cores on a
bounded sample of the same solve.

Other workloads: --workload powerlaw (8M x 4M Zipf bipartite, FP32 basis +
FP64 polish), --workload batch (256 uneven problems, batched vs per-problem
loop), --workload complex (ZLANSVD_IRL on a 1M x 500k complex matrix),
--workload spmv-probe (SpMV time by column pattern).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
# The reference's DCGS keeps an automatic work vector of the basis length on
# every OpenMP thread's stack (double/dreorth.F:132): 16 MB at 2M rows.
os.environ.setdefault("OMP_STACKSIZE", "512M")

from propack_amd import api, synth  # noqa: E402

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64 MFMA

M, N, LAM, DECAY, SEED = 2_000_000, 1_000_000, 16.0, 0.6, 2024
NEIG, DIM, TOL = 50, 120, 1e-10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_spmv_poisson_decay_2Mx1M.md")


def measured_traffic(kclass, panels, lanes, sort=(0, 0)):
    """HBM bytes per operator application of kernel class kclass, as
    measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/ (calibrated against a plain copy) for the SpMV layout in use
    (column panels, lanes-per-row shifts, sorted-layout block caps; a line
    without caps is the row-block layout).  The lane shift only regroups a
    row's additions inside a workgroup (the summary measures four shift
    pairs per layout within 1.5 % of each other), so without a line for
    these shifts the same panels and caps at other shifts answer; None if
    the layout was not measured."""
    want = (*tuple(panels), *tuple(lanes), *tuple(sort))
    same_bytes = None
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                w = line.split()
                if len(w) not in (7, 9) or w[0] != "traffic_bytes_per_application" \
                        or w[1] != kclass:
                    continue
                got = tuple(int(x) for x in w[3:] + ["0", "0"][: 9 - len(w)])
                if got == want:
                    return float(w[2])
                if same_bytes is None and got[:2] == want[:2] and got[4:] == want[4:]:
                    same_bytes = float(w[2])
    except OSError:
        pass
    return same_bytes


ZTRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_zspmv_complex_poisson_decay_1Mx500k.md")


def measured_ztraffic(kclass, panels):
    """HBM bytes per application of the complex SpMV class kclass for the
    operator's column panels (A, A^H), from the FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/; None if that layout was not measured."""
    try:
        with open(ZTRAFFIC_PROFILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 5 and w[0] == "ztraffic_bytes_per_application" and \
                        w[1] == kclass and (int(w[3]), int(w[4])) == tuple(panels):
                    return float(w[2])
    except OSError:
        pass
    return None


def residuals(op, r):
    """max_i ||A v_i - s_i u_i|| / s_1 over the returned triplets."""
    worst = 0.0
    for i in range(r.k):
        av = op.spmv(np.ascontiguousarray(r.V[:, i]), "n")
        worst = max(worst, float(np.linalg.norm(av - r.sigma[i] * r.U[:, i])))
    return worst / float(r.sigma[0])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(A, u0, gpu_nopx):
    """Reference PROPACK (oracle/_ref, OpenMP) on a bounded sample: one
    Lanczos pass of DLANSVD_IRL (maxiter = 1: 120 steps, the shifted
    restart and its basis GEMMs), scaled to the full solve by the ratio of
    operator applications (nopx) of the device run.  The median of three
    runs on all host cores is the value; one run on a single thread is
    reported beside it, with the reference's own counters."""
    try:
        from oracle import refpy
    except Exception as exc:  # pragma: no cover
        return {"error": f"oracle unavailable: {exc}"}
    if not refpy.available("d"):
        return {"error": "oracle/_ref not built"}
    cores = os.cpu_count() or 1
    AT = A.transpose()
    out = {}

    def run(threads):  # on a thread with a stack large enough for the reference
        refpy.set_threads(threads)  # OpenMP's thread count is per calling thread
        t0 = time.perf_counter()
        out["r"] = refpy.lansvd_irl(A, NEIG, DIM, u0=u0, tol=TOL, maxiter=1, cgs=1, elr=1,
                                    AT=AT, jobu="y", jobv="y")
        out["t"] = time.perf_counter() - t0

    threading.stack_size(1 << 30)

    def timed(threads):
        th = threading.Thread(target=run, args=(threads,))
        th.start()
        th.join()
        return out["r"], out["t"]

    runs = [timed(cores) for _ in range(3)]
    ts = sorted(t for _, t in runs)
    r, t = runs[0][0], ts[1]
    r1, t1 = timed(1)
    nopx = r.counters["nopx"]
    keep = ("nopx", "nreorth", "ndot", "nrestart", "nbsvd")
    return {"value": t * gpu_nopx / nopx, "unit": "s", "cores": cores, "kind": "reference",
            "sample": f"DLANSVD_IRL maxiter=1 (singular vectors on) on the same matrix, "
                      f"{nopx} operator applications, {r.k} converged, median of 3 runs "
                      f"on {cores} threads {t:.2f} s wall; scaled by the device run's "
                      f"{gpu_nopx} applications",
            "sample_s": t, "sample_s_runs": ts, "sample_nopx": nopx, "info": r.info,
            "counters": {k: r.counters[k] for k in keep},
            "one_thread": {"value": t1 * gpu_nopx / r1.counters["nopx"], "sample_s": t1,
                           "counters": {k: r1.counters[k] for k in keep}},
            "cpu_model": _cpu_model()}


def copy_calibration():
    """A plain device-to-device copy of exactly 2^30 bytes, five times: the
    HBM counters of this kernel give bytes-counted / bytes-moved."""
    import torch
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 5 x {src.numel() * 8} bytes")


def _timed(f, steps, sync):
    f()  # warm-up
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        out = f()
        sync()
        ts.append(time.perf_counter() - t0)
    return out, float(np.mean(ts))


def workload_powerlaw(args):
    """8,000,000 x 4,000,000 power-law bipartite adjacency (row degrees
    ~ d^-2.1 on [1, 1e5], each row's columns distinct and Zipf(1.8)-ranked
    over a seeded permutation, values 1.0 exact in FP32): k=64, dim=160,
    tol=1e-8.  FP64
    solve, FP32-basis solve + FP64 polish, and the FP32 phase alone."""
    k, dim, tol = 64, 160, 1e-8
    t0 = time.perf_counter()
    A = synth.powerlaw_bipartite(8_000_000, 4_000_000, seed=8000, dtype=np.float64)
    import torch
    # the start vector is an input: resident in HBM before the timed region
    u0 = torch.from_numpy(synth.start_vector(A.m, 8001)).to("cuda")
    log(f"[bench] power-law matrix {A.m}x{A.n} nnz={A.nnz} in {time.perf_counter() - t0:.1f}s")
    A32 = A.astype(np.float32)
    tu64 = None if args.no_tune else api.autotune(A)
    tu32 = None if args.no_tune else api.autotune(A32, prec=api.PK_F32)
    if tu64:
        log(f"[bench] SpMV autotune: FP64 panels {tu64.panels} lanes {tu64.lanes} sort "
            f"{tu64.sort}, FP32 panels {tu32.panels} lanes {tu32.lanes} sort {tu32.sort}")
    op = api.CsrOperator(A, panels=tu64.panels if tu64 else None, lanes=tu64.lanes if tu64 else None,
                         sort=tu64.sort if tu64 else None)
    eng = api.Engine(op, kmax=dim)
    op32 = api.CsrOperator(A32, prec=api.PK_F32, panels=tu32.panels if tu32 else None,
                           lanes=tu32.lanes if tu32 else None, sort=tu32.sort if tu32 else None)
    eng32 = api.Engine(op32, kmax=dim)
    sync = lambda: api.lib().pk_lbd_sync(eng.handle)  # noqa: E731

    def fp64():
        eng.put_u1(u0)
        return api.lansvd_irl_dev(eng, k, dim, tol=tol, fetch=False)
    r64, t64 = _timed(fp64, args.steps, sync)
    log(f"[bench] FP64: {t64:.3f}s k={r64.k} info={r64.info} {r64.counters}")

    def mixed():
        eng32.put_u1(u0)
        return api.lansvd_irl_mixed(eng, eng32, k, dim, tol=tol, fetch=False)
    (rm, f32), tm = _timed(mixed, args.steps, sync)
    e32 = 2.0 ** -24
    eng32.set_rounding(1)

    def fp32():
        eng32.put_u1(u0)
        return api.lansvd_irl_dev(eng32, k, dim, tol=max(tol, 16 * e32), delta=np.sqrt(e32),
                                  eta=e32 ** 0.75, fetch=False)
    r32, t32 = _timed(fp32, args.steps, lambda: api.lib().pk_lbd_sync(eng32.handle))
    kk = min(r64.k, f32["k"])
    err32 = np.abs(f32["sigma"][:kk] - r64.sigma[:kk]) / r64.sigma[:kk]
    errp = np.abs(rm.sigma[:kk] - r64.sigma[:kk]) / r64.sigma[:kk]
    resid = None
    if not args.no_check:
        rr = api.SvdResult(rm.sigma, rm.bnd, eng.get("u", rm.k), eng.get("v", rm.k), rm.info, rm.k)
        resid = residuals(op, rr)
    out = {"metric": "time-to-64 triplets at tol 1e-8, power-law 8Mx4M: FP32-basis + FP64 polish "
                     "vs FP64",
           "value": tm, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": tm * 1e3, "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "f32 storage / f64 accumulation + f64 polish",
           "data": "synthetic (seeded generator, matrix resident in HBM)",
           "config": {"workload": "DLANSVD_IRL k=64 dim=160 tol=1e-8 on powerlaw_bipartite "
                                  "8000000x4000000 (degrees d^-2.1, Zipf(1.8) columns)",
                      "m": A.m, "n": A.n, "nnz": A.nnz, "nnz_note": "each row's columns drawn "
                      "distinct (successive Zipf sampling)"},
           "fp64": {"s": t64, "converged": r64.k, "info": r64.info, "nopx": r64.counters["nopx"],
                    "nrestart": r64.counters["nrestart"]},
           "fp32_phase": {"s": t32, "converged": r32.k, "info": r32.info,
                          "nopx": r32.counters["nopx"],
                          "rel_sigma_err_vs_fp64_per_triplet": [float(x) for x in err32]},
           "polished": {"s": tm, "converged": rm.k, "info": rm.info,
                        "max_rel_sigma_err_vs_fp64": float(errp.max()) if kk else None,
                        "max_rel_residual": resid},
           "speedup_polished_vs_fp64": t64 / tm, "speedup_fp32_phase_vs_fp64": t64 / t32}
    print(json.dumps(out))


def workload_c3(args):
    """The FP32-basis form of the 1850 x 712 non-restarted solve (DLANSVD,
    k=10, kmax=300, tol=1e-12; 8636 uniform draws with N(0,1) values,
    duplicates summed): per-triplet sigma error and FP64 residual of the
    FP32 phase against the FP64 solve, the polished result, and the times.
    The matrix is cache-resident: launch latency, not HBM, sets the times."""
    k, kmax, tol = 10, 300, 1e-12
    A = synth.random_uniform()
    import torch
    u0 = torch.from_numpy(synth.start_vector(A.m, 1)).to("cuda")
    e64 = api.Engine(api.CsrOperator(A), kmax=kmax)
    e32 = api.Engine(api.CsrOperator(A, prec=api.PK_F32), kmax=kmax)
    sync = lambda: api.lib().pk_lbd_sync(e64.handle)  # noqa: E731

    def fp64():
        e64.put_u1(u0)
        return api.lansvd_dev(e64, k, kmax, tol=tol, fetch=False)
    r64, t64 = _timed(fp64, args.steps, sync)

    def mixed():
        e32.put_u1(u0)
        return api.lansvd_mixed(e64, e32, k, kmax, tol=tol, fetch=False)
    (rm, f32), tm = _timed(mixed, args.steps, sync)
    # the FP32 phase alone, with its right vectors, for the FP64 residuals
    e32.set_rounding(1)
    e32.put_u1(u0)
    e32_ = 2.0 ** -24
    r32 = api.lansvd_dev(e32, k, kmax, tol=16 * e32_, delta=np.sqrt(e32_), eta=e32_ ** 0.75,
                         fetch=False)
    kk = min(r64.k, r32.k)
    res32 = e64.residuals(e32, r32.sigma[:kk]) if kk else np.zeros((0, 2))
    s1 = r64.sigma[0] if r64.k else 1.0
    out = {"metric": "FP32-basis form of DLANSVD k=10 kmax=300 tol=1e-12 (1850x712): time and "
                     "per-triplet accuracy vs FP64",
           "value": tm, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": tm * 1e3, "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "f32 storage / f64 accumulation + f64 polish",
           "data": "synthetic (seeded generator, matrix resident in HBM)",
           "config": {"workload": "DLANSVD k=10 kmax=300 tol=1e-12 on random_uniform 1850x712 "
                                  "(8636 draws)", "m": A.m, "n": A.n, "nnz": A.nnz},
           "fp64": {"s": t64, "k": r64.k, "info": r64.info, "nopx": r64.counters["nopx"]},
           "fp32_phase": {"k": r32.k, "info": r32.info, "nopx": r32.counters["nopx"],
                          "rel_sigma_err_vs_fp64_per_triplet":
                              [float(x) for x in np.abs(r32.sigma[:kk] - r64.sigma[:kk])
                               / r64.sigma[:kk]],
                          "fp64_rel_residual_per_triplet":
                              [float(x) for x in res32[:, 0] / s1]},
           "polished": {"s": tm, "k": rm.k, "info": rm.info,
                        "max_rel_sigma_err_vs_fp64":
                            float(np.max(np.abs(rm.sigma[:kk] - r64.sigma[:kk]) / r64.sigma[:kk]))
                            if kk else None},
           "speedup_vs_fp64": t64 / tm}
    print(json.dumps(out))


def workload_complex(args):
    """ZLANSVD_IRL (complex16/zlansvd_irl.F) on a complex 1,000,000 x 500,000
    CSR matrix: the poisson_decay pattern (Poisson(16) distinct columns per
    row, |a_ij| = |N(0,1)| (1 + j)^-0.6) with a seeded unit phase per value;
    k=50, dim=120, tol=1e-10, complex FP64 bases, start vector resident in
    HBM.  Live timing of every kernel class in one more solve; the complex
    SpMV's achieved bandwidth against HBM peak; residuals of three triplets
    on the host; the reference ZLANSVD_IRL (oracle/_ref) on the host cores on
    one Lanczos pass (maxiter = 1), scaled by operator applications."""
    k, dim, tol = 50, 120, 1e-10
    m, n, seed = 1_000_000, 500_000, 2025
    t0 = time.perf_counter()
    A = synth.complex_poisson_decay(m, n, seed=seed)
    u0h = synth.start_vector(m, seed + 2, np.complex128)
    import torch
    u0 = torch.from_numpy(u0h).to("cuda")
    log(f"[bench] complex matrix {A.m}x{A.n} nnz={A.nnz} in {time.perf_counter() - t0:.1f}s")
    op = api.CsrOperator(A, api.PK_C128)
    eng = api.Engine(op, kmax=dim)
    sync = lambda: api.lib().pk_lbd_sync(eng.handle)  # noqa: E731

    def solve():
        eng.put_u1(u0)
        return api.lansvd_irl_dev(eng, k, dim, tol=tol, jobu="y", jobv="y", fetch=False)
    r, t = _timed(solve, args.steps, sync)
    log(f"[bench] complex solve {t:.4f} s, k={r.k} info={r.info} {r.counters}")
    eng.profile(True)
    eng.put_u1(u0)
    rp = api.lansvd_irl_dev(eng, k, dim, tol=tol, jobu="y", jobv="y", fetch=False)
    sync()
    prof = eng.profile_get()
    eng.profile(False)
    kern = {c: dict(v, GBps=(v["bytes"] / (v["ms"] * 1e6) if v["ms"] else None))
            for c, v in prof.items()}
    # the dominant kernel: the SpMV class with more time
    dom = max(("spmv_n", "spmv_t"), key=lambda c: prof[c]["ms"])
    pd = prof[dom]
    achieved = pd["bytes"] / (pd["ms"] * 1e6) if pd["ms"] else 0.0
    # residuals of three triplets on the host
    nres = min(3, r.k)
    resid = None
    if nres and not args.no_check:
        U = eng.get("u", nres)
        V = eng.get("v", nres)
        res = [np.linalg.norm(A.matvec(V[:, i]) - r.sigma[i] * U[:, i]) for i in range(nres)]
        resid = float(max(res) / r.sigma[0])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_complex(A, u0h, r.counters["nopx"], k, dim, tol)
    out = {"metric": "ZLANSVD_IRL time to 50 triplets at resid<=1e-10 (complex FP64 basis)",
           "value": t, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "c128",
           "data": "synthetic (seeded generator, matrix resident in HBM)",
           "config": {"workload": f"ZLANSVD_IRL k={k} dim={dim} tol={tol} on "
                                  f"complex_poisson_decay {m}x{n} (Poisson(16) per row, "
                                  "decay 0.6, unit phases)",
                      "m": A.m, "n": A.n, "nnz": A.nnz, "spmv_panels": list(op.panels()),
                      "converged": r.k, "info": r.info,
                      "nopx": r.counters["nopx"], "nreorth": r.counters["nreorth"],
                      "nrestart": r.counters["nrestart"], "max_rel_residual_3": resid},
           "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": measured_ztraffic(dom, op.panels()),
                        "avg_launch_ms": pd["ms"] / max(pd["launches"], 1),
                        "bytes_per_launch": pd["bytes"] / max(pd["launches"], 1),
                        "timed_in": "one profiled solve after the timed ones"},
           "kernels": kern,
           "profiled_solve_nopx": rp.counters["nopx"],
           "cpu_baseline": cpu}
    print(json.dumps(out))


def cpu_baseline_complex(A, u0, gpu_nopx, k, dim, tol):
    """Reference ZLANSVD_IRL (oracle/_ref) on one Lanczos pass (maxiter = 1),
    scaled by the device run's operator applications; median of 3 runs on
    all host cores (a one-thread run of this sample takes over a minute and
    is left out).  The reference runs its
    modified Gram-Schmidt (ioption(1) = 0): its complex classical
    Gram-Schmidt races on many threads (complex16/zreorth.F:140-201)."""
    try:
        from oracle import refpy
    except Exception as exc:  # pragma: no cover
        return {"error": f"oracle unavailable: {exc}"}
    if not refpy.available("z"):
        return {"error": "oracle/_ref not built"}
    cores = os.cpu_count() or 1
    AT = A.transpose()
    out = {}

    def run(threads):
        refpy.set_threads(threads, "z")
        t0 = time.perf_counter()
        out["r"] = refpy.lansvd_irl(A, k, dim, u0=u0, tol=tol, maxiter=1, cgs=0, elr=1, AT=AT,
                                    jobu="y", jobv="y", prec="z")
        out["t"] = time.perf_counter() - t0

    threading.stack_size(1 << 30)

    def timed(threads):
        th = threading.Thread(target=run, args=(threads,))
        th.start()
        th.join()
        return out["r"], out["t"]

    runs = [timed(cores) for _ in range(3)]
    ts = sorted(t for _, t in runs)
    r, t = runs[0][0], ts[1]
    nopx = r.counters["nopx"]
    keep = ("nopx", "nreorth", "ndot", "nrestart", "nbsvd")
    return {"value": t * gpu_nopx / nopx, "unit": "s", "cores": cores, "kind": "reference",
            "sample": f"ZLANSVD_IRL maxiter=1 (MGS, singular vectors on) on the same matrix, "
                      f"{nopx} operator applications, median of 3 runs on {cores} threads "
                      f"{t:.2f} s wall; scaled by the device run's {gpu_nopx} applications",
            "sample_s": t, "sample_s_runs": ts, "sample_nopx": nopx, "info": r.info,
            "counters": {c: r.counters[c] for c in keep},
            "cpu_model": _cpu_model()}


def workload_spmv_probe(args):
    """Where the SpMV time goes: A x on the bench matrix's rows (same row
    pointers, values and nonzero count) with three column patterns: every
    row's k-th nonzero in column k (all gathers hit a few cache lines: L1),
    the original columns folded into a 64K-column window (512 KiB of x:
    L2-resident), and the original uniform columns over 1M (the bench
    matrix).  Each pattern's best time per application over the autotuner's
    layouts; the first two bound what the gathers' misses cost."""
    A = synth.poisson_decay(M, N, LAM, DECAY, seed=SEED)
    rp = A.rowptr.astype(np.int64)
    lens = np.diff(rp)
    pos = np.arange(A.nnz, dtype=np.int64) - np.repeat(rp[:-1], lens)
    pats = {"l1_same_lines": pos.astype(np.int32),
            "l2_window_64k": (A.col % 65536).astype(np.int32),
            "uniform_1m": A.col}
    res = {}
    alg = A.nnz * 12.0 + (A.m + 1) * 4.0 + A.n * 8.0 + 2 * A.m * 8.0
    for name, cols in pats.items():
        B = synth.Csr(A.m, A.n, A.rowptr, np.ascontiguousarray(cols), A.val)
        t = api.autotune(B)
        res[name] = {"us_per_application": round(t.us[0], 1), "panels": t.panels[0],
                     "lane_shift": t.lanes[0], "sort_cap": t.sort[0],
                     "algorithmic_GBps": round(alg / (t.us[0] * 1e-6) / 1e9, 1)}
        log(f"[bench] spmv probe {name}: {res[name]}")
    out = {"metric": "SpMV A x time per application by column pattern (2Mx1M, 32M nnz, FP64)",
           "value": res["uniform_1m"]["us_per_application"], "unit": "us", "n_gpus": 1,
           "steps": 5, "warmup": 1, "higher_is_better": False, "dtype": "f64",
           "data": "synthetic (rows, values of the bench matrix; columns per pattern)",
           "config": {"workload": "SpMV column-pattern probe", "nnz": A.nnz,
                      "algorithmic_bytes": alg}, "patterns": res}
    print(json.dumps(out))


def workload_batch(args):
    """The batched workload (synth.batch_params: m log-uniform in [5e3, 5e5],
    n = r m, Poisson(lam ~ U[4, 48]) nonzeros per row, decay p ~ U[0.3, 1]):
    k=16, dim=48, tol=1e-8; problems/s of the batched solve against a
    per-problem loop on the same engines, in the FP32-basis form (the
    reported value) and in FP64, with each form's worst residual
    ||A v - s u|| / s_1 formed on the host after the timed runs."""
    import concurrent.futures as cf
    import gc

    import scipy.sparse as sp
    import torch
    t0 = time.perf_counter()
    # column-sorted SpMV blocks of 8192 nonzeros for every problem, a quarter
    # of the built-in lanes per row: the fastest fixed setting on this
    # workload (measured problems/s: row-block layout 417; caps 4096, 8192,
    # 16384 at the built-in lanes 462, 532, 435; cap 8192 with lane shifts
    # 1, 0, -1, -2: 474, 529, 546-552, 558)
    lanes = tuple(int(x) for x in args.lanes.split(",")) if args.lanes else (-2, -2)
    sort = tuple(int(x) for x in args.sort.split(",")) if args.sort else (8192, 8192)
    params = synth.batch_params(256)[: args.count]
    with cf.ThreadPoolExecutor(8) as ex:
        mats = list(ex.map(lambda prm: synth.batch_problem(prm, dtype=np.float32), params))
    nnz = sum(A.nnz for A in mats)
    log(f"[bench] {len(mats)} problems, {nnz / 1e6:.1f}M nnz generated in "
        f"{time.perf_counter() - t0:.1f}s")
    forms = {}
    for basis in ("fp32", "fp64"):
        prec = api.PK_F32 if basis == "fp32" else api.PK_F64
        t0 = time.perf_counter()
        probs = []
        for prm, A in zip(params, mats):
            Ap = A if prec == api.PK_F32 else A.astype(np.float64)
            if args.tune_each:
                t = api.autotune(Ap, prec=prec)
                op = api.CsrOperator(Ap, prec=prec, panels=t.panels, lanes=t.lanes, sort=t.sort)
            else:
                op = api.CsrOperator(Ap, prec=prec, lanes=lanes, sort=sort)
            e = api.Engine(op, kmax=48, prec=prec)
            if prec == api.PK_F32:
                e.set_rounding(1)
            # start vectors are inputs: resident in HBM before the timed region
            u0 = torch.from_numpy(synth.start_vector(A.m, prm["seed"] + 7)).to("cuda")
            probs.append((op, e, u0, Ap))
            if len(probs) % 64 == 0:
                log(f"[bench] {basis}: {len(probs)} problems set up "
                    f"({time.perf_counter() - t0:.1f}s)")
        sync = lambda: api.lib().pk_lbd_sync(probs[0][1].handle)  # noqa: E731

        def seq():
            out = []
            for op, e, u0, _ in probs:
                e.put_u1(u0)
                out.append(api.lansvd_irl_dev(e, 16, 48, tol=1e-8, jobu="n", jobv="n",
                                              fetch=False))
            return out
        # --split G: G batches over interleaved problems (largest first, dealt
        # round-robin), each run from a host thread of its own, so that one
        # batch's host work between restarts overlaps the others' GPU work
        order = sorted(range(len(probs)), key=lambda i: -probs[i][3].nnz)
        groups = [order[g::args.split] for g in range(args.split)]
        bs = [api.Batch([probs[i][1] for i in grp]) for grp in groups]
        for b in bs:
            b.set_graph(not args.no_graph)

        def bat(fetch=False):
            for op, e, u0, _ in probs:
                e.put_u1(u0)
            out = [None] * len(probs)

            def run(g):
                res = api.lansvd_irl_batch(bs[g], 16, 48, tol=1e-8, jobu="y" if fetch else "n",
                                           jobv="y" if fetch else "n", fetch=fetch)
                for i, r in zip(groups[g], res):
                    out[i] = r
            if len(bs) == 1:
                run(0)
            else:
                ths = [threading.Thread(target=run, args=(g,)) for g in range(len(bs))]
                for t in ths:
                    t.start()
                for t in ths:
                    t.join()
            return out
        rs, ts = _timed(seq, args.steps, sync)
        rb, tb = _timed(bat, args.steps, sync)
        same = all(np.array_equal(x.sigma, y.sigma) for x, y in zip(rs, rb))
        # the same batched solve with its bidiagonal SVDs on the device
        for b in bs:
            b.set_bsvd(True)
        rd, td = _timed(bat, args.steps, sync)
        for b in bs:
            b.set_bsvd(False)
        dev = {"problems_per_s": len(probs) / td,
               "converged": sum(r.k for r in rd),
               "max_rel_sigma_diff_vs_host_dbdsqr": float(max(
                   np.abs(x.sigma - y.sigma).max() / x.sigma[0] for x, y in zip(rb, rd)
                   if x.k == y.k and x.k > 0)),
               "same_converged_counts": all(x.k == y.k for x, y in zip(rb, rd))}
        worst = 0.0
        for r, (op, e, u0, A) in zip(bat(fetch=True), probs):
            M = sp.csr_array((A.val.astype(np.float64), A.col, A.rowptr), shape=(A.m, A.n))
            res = np.linalg.norm(M @ r.V.astype(np.float64) - r.U.astype(np.float64) * r.sigma,
                                 axis=0).max() / r.sigma[0]
            worst = max(worst, float(res))
        forms[basis] = {"problems_per_s": len(probs) / tb, "ms_per_step": tb * 1e3,
                        "sequential_problems_per_s": len(probs) / ts,
                        "speedup_vs_sequential": ts / tb,
                        "converged": sum(r.k for r in rb), "requested": 16 * len(probs),
                        "bit_identical_to_sequential": same,
                        "max_rel_residual": worst, "device_bsvd": dev}
        log(f"[bench] {basis}: {forms[basis]}")
        for b in bs:
            b.close()
        del probs, bs, rs, rb
        gc.collect()
    f32 = forms["fp32"]
    out = {"metric": "batched truncated SVDs per second (k=16, dim=48, tol=1e-8, FP32 basis)",
           "value": f32["problems_per_s"], "unit": "problems/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 1, "ms_per_step": f32["ms_per_step"],
           "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32 storage / f64 accumulation",
           "data": "synthetic (seeded generator, matrices resident in HBM)",
           "config": {"workload": f"{len(mats)} problems of batch_params(256)", "nnz": nnz,
                      "spmv_layout": ("autotuned per problem" if args.tune_each else
                                      {"sort_caps": list(sort), "lane_shifts": list(lanes)}),
                      "hip_graphs": not args.no_graph, "batches": args.split},
           "sequential_problems_per_s": f32["sequential_problems_per_s"],
           "speedup_vs_sequential": f32["speedup_vs_sequential"],
           "converged": f32["converged"], "requested": f32["requested"],
           "bit_identical_to_sequential": f32["bit_identical_to_sequential"],
           "max_rel_residual": f32["max_rel_residual"],
           "device_bsvd": f32["device_bsvd"],
           "fp64_basis": forms["fp64"]}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-gemm", action="store_true",
                    help="skip the restart-GEMM timing at the headline shape")
    ap.add_argument("--no-tune", action="store_true",
                    help="built-in SpMV column panels instead of the autotuner's")
    ap.add_argument("--panels", default=None,
                    help="explicit SpMV column panels 'A,At' (no autotuning)")
    ap.add_argument("--lanes", default=None,
                    help="lanes-per-row shifts 'A,At' with --panels (default 0,0), or of "
                         "every problem of --workload batch (default -2,-2)")
    ap.add_argument("--sort", default=None,
                    help="column-sorted SpMV layout block caps 'A,At' with --panels "
                         "(default 0,0: row-block layout), or of every problem of --workload "
                         "batch (default 8192,8192)")
    ap.add_argument("--workload", choices=["c2", "c3", "powerlaw", "batch", "complex",
                                           "spmv-probe"],
                    default="c2",
                    help="c2 (default): the headline solve; powerlaw: 8M x 4M Zipf "
                         "bipartite, k=64, dim=160, tol=1e-8, FP32 basis + FP64 polish vs "
                         "FP64; batch: the 256-problem batched workload vs a per-problem loop; "
                         "c3: the FP32-basis form of the 1850x712 DLANSVD solve; complex: "
                         "ZLANSVD_IRL on a 1M x 500k complex matrix")
    ap.add_argument("--count", type=int, default=256, help="problems of --workload batch")
    ap.add_argument("--no-graph", action="store_true",
                    help="--workload batch: enqueue the batched step chains directly instead of "
                         "replaying captured HIP graphs")
    ap.add_argument("--split", type=int, default=1,
                    help="--workload batch: run the problems as this many batches, each from a "
                         "host thread of its own")
    ap.add_argument("--tune-each", action="store_true",
                    help="--workload batch: autotune every problem's SpMV layout")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the reference "
                         "kernel for scaling FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU layout is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "powerlaw":
        workload_powerlaw(args)
        return
    if args.workload == "c3":
        workload_c3(args)
        return
    if args.workload == "batch":
        workload_batch(args)
        return
    if args.workload == "spmv-probe":
        workload_spmv_probe(args)
        return
    if args.workload == "complex":
        workload_complex(args)
        return

    t0 = time.perf_counter()
    A = synth.poisson_decay(M, N, LAM, DECAY, seed=SEED)
    u0 = synth.start_vector(M, SEED + 1)
    log(f"[bench] matrix {A.m}x{A.n} nnz={A.nnz} generated in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    panels, lanes, sort, tuned = None, None, None, None
    if args.panels:
        panels = tuple(int(x) for x in args.panels.split(","))
        lanes = tuple(int(x) for x in args.lanes.split(",")) if args.lanes else None
        sort = tuple(int(x) for x in args.sort.split(",")) if args.sort else None
    elif not args.no_tune:
        tuned = api.autotune(A)
        panels, lanes, sort = tuned.panels, tuned.lanes, tuned.sort
        log(f"[bench] SpMV autotune: panels A={panels[0]} A^T={panels[1]}, lane shifts "
            f"A={lanes[0]} A^T={lanes[1]}, sorted-layout caps A={sort[0]} A^T={sort[1]} "
            f"({tuned.us[0]:.1f} / {tuned.us[1]:.1f} us per application) in "
            f"{time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    op = api.CsrOperator(A, panels=panels, lanes=lanes, sort=sort)
    eng = api.Engine(op, kmax=DIM)
    log(f"[bench] device operator + A^T + engine in {time.perf_counter() - t0:.1f}s")

    # the start vector is an input: resident in HBM before the timed region
    import torch
    u0_dev = torch.from_numpy(u0).to("cuda")

    def solve():
        eng.put_u1(u0_dev)
        return api.lansvd_irl_dev(eng, NEIG, DIM, tol=TOL, fetch=False)

    for w in range(args.warmup):
        t0 = time.perf_counter()
        r = solve()
        log(f"[bench] warmup {w}: {time.perf_counter() - t0:.4f}s info={r.info} k={r.k} "
            f"{r.counters}")

    # Live timing inside the timed solves covers the SpMV classes only (the
    # roofline kernel): every timed region adds two marker packets to the
    # stream.  One more solve, outside the timed ones, times every class.
    eng.profile(True, classes=api.Engine.PROF_SPMV)
    eng.reset_stats()
    times = []
    api.lib().pk_lbd_sync(eng.handle)
    for s in range(args.steps):
        t0 = time.perf_counter()
        r = solve()
        api.lib().pk_lbd_sync(eng.handle)
        times.append(time.perf_counter() - t0)
        log(f"[bench] step {s}: {times[-1]:.4f}s info={r.info} k={r.k} {r.counters}")
    prof_timed = eng.profile_get()
    stats = eng.stats()
    eng.profile(True)
    t0 = time.perf_counter()
    solve()
    api.lib().pk_lbd_sync(eng.handle)
    log(f"[bench] solve with every kernel class timed: {time.perf_counter() - t0:.4f}s")
    prof_all = eng.profile_get()
    eng.profile(False)
    ms_per_step = 1000.0 * float(np.mean(times))

    # dominant kernel class by device time (all classes timed); its figures
    # come from the timed solves when it was timed there
    dom = max(prof_all, key=lambda c: prof_all[c]["ms"])
    p = prof_timed[dom] if prof_timed[dom]["launches"] else prof_all[dom]
    per_launch_ms = p["ms"] / max(p["launches"], 1)
    per_launch_bytes = p["bytes"] / max(p["launches"], 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    kernels = {c: {"launches": v["launches"], "ms_total": round(v["ms"], 3),
                   "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None,
                   "bytes_per_launch": v["bytes"] / v["launches"] if v["launches"] else 0}
               for c, v in prof_all.items()}
    g = prof_all["gemm"]
    if g["ms"] > 0:
        kernels["gemm"]["TFLOPs"] = round(g["flops"] / (g["ms"] * 1e-3) / 1e12, 2)
        kernels["gemm"]["mfma_util"] = round(g["flops"] / (g["ms"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)

    # SpMV and reorthogonalization together (the two HBM-bound classes):
    # algorithmic bytes over device time
    cls = ("spmv_n", "spmv_t", "gemvt", "gemv")
    cb = sum(prof_all[c]["bytes"] for c in cls)
    cms = sum(prof_all[c]["ms"] for c in cls)
    combined = ({"GBps": round(cb / (cms * 1e-3) / 1e9, 1),
                 "frac": round(cb / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "ms": round(cms, 3), "bytes": cb} if cms > 0 else None)
    log(f"[bench] kernels: {json.dumps(kernels)}")
    log(f"[bench] engine stats: {stats}")

    resid = None
    if not args.no_check:
        rr = api.SvdResult(r.sigma, r.bnd, eng.get("u", r.k), eng.get("v", r.k), r.info, r.k)
        resid = residuals(op, rr)
        log(f"[bench] max ||A v - s u|| / s_1 = {resid:.3e}")

    # FP32-basis form with FP64 polish on the same matrix and start vector
    mixed = None
    if not args.no_mixed:
        A32 = A.astype(np.float32)
        t32 = None if args.no_tune else api.autotune(A32, prec=api.PK_F32)
        op32 = api.CsrOperator(A32, prec=api.PK_F32, panels=t32.panels if t32 else None,
                               lanes=t32.lanes if t32 else None, sort=t32.sort if t32 else None)
        eng32 = api.Engine(op32, kmax=DIM)
        u32 = torch.from_numpy(u0.astype(np.float32).astype(np.float64)).to("cuda")

        def solve_mixed():
            eng32.put_u1(u32)
            return api.lansvd_irl_mixed(eng, eng32, NEIG, DIM, tol=TOL, fetch=False)

        rm, f32 = solve_mixed()  # warm-up
        tm = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            rm, f32 = solve_mixed()
            api.lib().pk_lbd_sync(eng.handle)
            tm.append(time.perf_counter() - t0)
        # the FP32 phase alone (the same solve the mixed driver runs first)
        eng32.set_rounding(1)
        e32 = 2.0 ** -24

        def solve_fp32():
            eng32.put_u1(u32)
            return api.lansvd_irl_dev(eng32, NEIG, DIM, tol=max(TOL, 16 * e32),
                                      delta=np.sqrt(e32), eta=e32 ** 0.75, fetch=False)

        solve_fp32()
        t32 = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            r32 = solve_fp32()
            api.lib().pk_lbd_sync(eng32.handle)
            t32.append(time.perf_counter() - t0)
        k32 = min(f32["k"], r.k)
        err32 = np.abs(f32["sigma"][:k32] - r.sigma[:k32]) / r.sigma[:k32]
        # the FP32 phase's triplets checked with the FP64 operator on the
        # device (pk_lbd_residuals): ||A v - s u|| / s_1 per triplet
        res32 = eng.residuals(eng32, r32.sigma[:r32.k])[:, 0] / r.sigma[0]
        errp = np.abs(rm.sigma[:r.k] - r.sigma[:r.k]) / r.sigma[:r.k] if rm.k >= r.k else None
        mixed = {"s": float(np.mean(tm)), "speedup_vs_fp64": float(ms_per_step / 1000.0 / np.mean(tm)),
                 "fp32_phase_converged": f32["k"], "fp32_phase_info": f32["info"],
                 "fp32_phase_max_rel_sigma_err_vs_fp64": float(err32.max()) if k32 else None,
                 "fp32_phase_rel_sigma_err_per_triplet": [float(x) for x in err32],
                 "polished_converged": rm.k,
                 "polished_max_rel_sigma_err_vs_fp64": float(errp.max()) if errp is not None else None,
                 "nopx_total": rm.counters["nopx"],
                 "polish_nopx": rm.counters["nopx"] - r32.counters["nopx"],
                 "fp32_phase_fp64_rel_residual_per_triplet": [float(x) for x in res32],
                 "fp32_phase_fp64_max_rel_residual": float(res32.max()),
                 "triplets_missing_tol_before_polish": int((res32 > TOL).sum()),
                 "fp32_phase_s": float(np.mean(t32)), "fp32_phase_nopx": r32.counters["nopx"],
                 "fp32_phase_speedup_vs_fp64": float(ms_per_step / 1000.0 / np.mean(t32))}
        log(f"[bench] FP32-basis + FP64 polish: {mixed['s']:.3f}s "
            f"(x{mixed['speedup_vs_fp64']:.2f} vs FP64), fp32 phase k={f32['k']} "
            f"max sigma err {mixed['fp32_phase_max_rel_sigma_err_vs_fp64']}; FP32 phase alone "
            f"{mixed['fp32_phase_s']:.3f}s (x{mixed['fp32_phase_speedup_vs_fp64']:.2f})")

    # The restart GEMM at the headline shape (DGEMM_OVWR_LEFT,
    # double/dlansvd_irl.F:387): U(:, 1:51) <- U(:, 1:121) Q, 2,000,000 x 121
    # times 121 x 51, FP64 MFMA; the solve above converges without a restart,
    # so it is timed here on its own (two untimed launches, then ten timed).
    restart_gemm = None
    if not args.no_gemm:
        Q = np.asfortranarray(np.random.default_rng(5).standard_normal((DIM + 1, NEIG + 1)))
        upd = lambda: api.lib().pk_lbd_update(  # noqa: E731
            eng.handle, api.C.c_char(b"u"), api.C.c_char(b"n"), api.C.c_int(DIM + 1),
            api.C.c_int(NEIG + 1), Q.ctypes.data_as(api.C.c_void_p), api.C.c_int(DIM + 1))
        for _ in range(2):
            upd()
        api.lib().pk_lbd_sync(eng.handle)
        eng.profile(True, classes=1 << 4)
        for _ in range(10):
            upd()
        pg = eng.profile_get()["gemm"]
        eng.profile(False)
        if pg["launches"]:
            ms = pg["ms"] / pg["launches"]
            tf = pg["flops"] / (pg["ms"] * 1e-3) / 1e12
            restart_gemm = {"shape": f"{A.m}x{DIM + 1} * {DIM + 1}x{NEIG + 1}",
                            "ms_per_launch": round(ms, 4), "TFLOPs": round(tf, 2),
                            "mfma_util": round(tf / FP64_PEAK_TFLOPS, 4),
                            "GBps": round(pg["bytes"] / (pg["ms"] * 1e-3) / 1e9, 1)}
            log(f"[bench] restart GEMM {restart_gemm}")

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (reference PROPACK, bounded sample) ...")
        cpu = cpu_baseline(A, u0, r.counters["nopx"])
        log(f"[bench] cpu baseline: {cpu}")

    out = {
        "metric": "time-to-50 triplets at resid<=1e-10 (FP64 basis); achieved HBM GB/s",
        "value": ms_per_step / 1000.0,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox generator, matrix resident in HBM)",
        "config": {"workload": "DLANSVD_IRL k=50 dim=120 tol=1e-10 on poisson_decay "
                               "2000000x1000000 (Poisson(16) per row, decay 0.6)",
                   "m": A.m, "n": A.n, "nnz": A.nnz, "k": NEIG, "dim": DIM, "tol": TOL,
                   "basis": "fp64", "converged": r.k, "info": r.info,
                   "nopx": r.counters["nopx"], "nreorth": r.counters["nreorth"],
                   "nrestart": r.counters["nrestart"], "host_syncs": stats["nsync"],
                   "spmv_panels": list(op.panels()), "spmv_lane_shifts": list(op.lanes()),
                   "spmv_sort_caps": list(op.sort()),
                   "autotuned": tuned is not None,
                   "max_rel_residual": resid},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, op.panels(), op.lanes(), op.sort()),
                     "avg_launch_ms": per_launch_ms, "bytes_per_launch": per_launch_bytes,
                     "timed_in": "the timed solves" if prof_timed[dom]["launches"] else
                                 "an extra solve with every class timed"},
        "kernels": kernels,
        "kernels_timed_in": "one extra solve with every kernel class timed (after the timed solves)",
        "spmv_plus_reorth": combined,
        "fp32_basis_form": mixed,
        "restart_gemm": restart_gemm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
