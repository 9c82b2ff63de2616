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
The matrix is resident in HBM (device CSR operator) before timing starts.

Also reported: the same solve's per-kernel live timings (HIP events on the
engine stream), the max relative residual of the 50 triplets, and the
reference PROPACK (OpenMP build, oracle/_ref) timed on the host cores on a
bounded sample of the same solve.

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


def measured_traffic(kclass):
    """HBM bytes per operator application of kernel class kclass, as
    measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/ (calibrated against a plain copy); None if absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 3 and w[0] == "traffic_bytes_per_application" and w[1] == kclass:
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


def cpu_baseline(A, u0, gpu_nopx):
    """Reference PROPACK (oracle/_ref, OpenMP) on a bounded sample: one
    Lanczos pass of DLANSVD_IRL (maxiter = 1: 120 steps, the shifted
    restart and its basis GEMMs), scaled to the full solve by the ratio of
    operator applications (nopx) of the device run."""
    try:
        from oracle import refpy
    except Exception as exc:  # pragma: no cover
        return {"error": f"oracle unavailable: {exc}"}
    if not refpy.available("d"):
        return {"error": "oracle/_ref not built"}
    cores = os.cpu_count() or 1
    refpy.set_threads(cores)
    AT = A.transpose()
    out = {}

    def run():  # on a thread with a stack large enough for the reference
        t0 = time.perf_counter()
        out["r"] = refpy.lansvd_irl(A, NEIG, DIM, u0=u0, tol=TOL, maxiter=1, cgs=1, elr=1,
                                    AT=AT, jobu="y", jobv="y")
        out["t"] = time.perf_counter() - t0

    threading.stack_size(1 << 30)
    th = threading.Thread(target=run)
    th.start()
    th.join()
    r, t = out["r"], out["t"]
    nopx = r.counters["nopx"]
    return {"value": t * gpu_nopx / nopx, "unit": "s", "cores": cores, "kind": "reference",
            "sample": f"DLANSVD_IRL maxiter=1 (singular vectors on) on the same matrix, "
                      f"{nopx} operator applications, {r.k} converged, {t:.2f} s wall; "
                      f"scaled by the device run's {gpu_nopx} applications",
            "sample_s": t, "sample_nopx": nopx}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the reference "
                         "kernel for scaling FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU layout is a design note only)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    A = synth.poisson_decay(M, N, LAM, DECAY, seed=SEED)
    u0 = synth.start_vector(M, SEED + 1)
    log(f"[bench] matrix {A.m}x{A.n} nnz={A.nnz} generated in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    op = api.CsrOperator(A)
    eng = api.Engine(op, kmax=DIM)
    log(f"[bench] device operator + A^T + engine in {time.perf_counter() - t0:.1f}s")

    def solve():
        eng.put_u1(u0)
        return api.lansvd_irl_dev(eng, NEIG, DIM, tol=TOL, fetch=False)

    for w in range(args.warmup):
        t0 = time.perf_counter()
        r = solve()
        log(f"[bench] warmup {w}: {time.perf_counter() - t0:.2f}s info={r.info} k={r.k} "
            f"{r.counters}")

    eng.profile(True)
    eng.reset_stats()
    times = []
    api.lib().pk_lbd_sync(eng.handle)
    for s in range(args.steps):
        t0 = time.perf_counter()
        r = solve()
        api.lib().pk_lbd_sync(eng.handle)
        times.append(time.perf_counter() - t0)
        log(f"[bench] step {s}: {times[-1]:.3f}s info={r.info} k={r.k} {r.counters}")
    prof = eng.profile_get()
    stats = eng.stats()
    eng.profile(False)
    ms_per_step = 1000.0 * float(np.mean(times))

    # dominant kernel class by accumulated device time
    dom = max(prof, key=lambda c: prof[c]["ms"])
    p = prof[dom]
    per_launch_ms = p["ms"] / max(p["launches"], 1)
    per_launch_bytes = p["bytes"] / max(p["launches"], 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    kernels = {c: {"launches": v["launches"], "ms_total": round(v["ms"], 3),
                   "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None,
                   "bytes_per_launch": v["bytes"] / v["launches"] if v["launches"] else 0}
               for c, v in prof.items()}

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
        op32 = api.CsrOperator(A.astype(np.float32), prec=api.PK_F32)
        eng32 = api.Engine(op32, kmax=DIM)
        u32 = u0.astype(np.float32).astype(np.float64)

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
        k32 = min(f32["k"], r.k)
        err32 = np.abs(f32["sigma"][:k32] - r.sigma[:k32]) / r.sigma[:k32]
        errp = np.abs(rm.sigma[:r.k] - r.sigma[:r.k]) / r.sigma[:r.k] if rm.k >= r.k else None
        mixed = {"s": float(np.mean(tm)), "speedup_vs_fp64": float(ms_per_step / 1000.0 / np.mean(tm)),
                 "fp32_phase_converged": f32["k"], "fp32_phase_info": f32["info"],
                 "fp32_phase_max_rel_sigma_err_vs_fp64": float(err32.max()) if k32 else None,
                 "fp32_phase_rel_sigma_err_per_triplet": [float(x) for x in err32],
                 "polished_converged": rm.k,
                 "polished_max_rel_sigma_err_vs_fp64": float(errp.max()) if errp is not None else None,
                 "nopx_total": rm.counters["nopx"]}
        log(f"[bench] FP32-basis + FP64 polish: {mixed['s']:.3f}s "
            f"(x{mixed['speedup_vs_fp64']:.2f} vs FP64), fp32 phase k={f32['k']} "
            f"max sigma err {mixed['fp32_phase_max_rel_sigma_err_vs_fp64']}")

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
                   "max_rel_residual": resid},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                     "avg_launch_ms": per_launch_ms, "bytes_per_launch": per_launch_bytes},
        "kernels": kernels,
        "fp32_basis_form": mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
