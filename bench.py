#!/usr/bin/env python3
"""Benchmark: time per right-hand side to a 1e-10 FP64 relative residual on a
512x512 lattice with a 3-level K-cycle DD-alpha-AMG preconditioner: the mixed-
precision batched solve (32 right-hand sides at once) next to the plain FP64
GPU solve (one right-hand side), on one MI355X.

Workload: quenched U(1) links from 500 Metropolis sweeps of the plaquette
action at beta = 3 (cold start, seeded), m0 = -0.20, parameters.dat rows
"0 8 8 10 4 4" and "1 4 4 10 4 4", Gaussian complex right-hand sides.

A step is one batched solve of the 32 right-hand sides.  Inputs are resident
in device memory before the timed region; set-up (test vectors, coarse
operators) runs before it and is reported separately.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time/RHS to 1e-10 residual, 512² 3-lvl: mixed+batched vs plain FP64 GPU"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--rhs", type=int, default=32)
    ap.add_argument("--sweeps", type=int, default=500)
    ap.add_argument("--cpu-iters", type=int, default=2, help="outer iterations of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    import amg_schwingermodel_amd as amgs

    nx = nt = a.nx
    rows = [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)]
    m0, beta = -0.20, 3.0
    t0 = time.time()
    U, plaq = amgs.gauge_metropolis(nx, nt, beta, a.sweeps, 20260512)
    log(f"gauge field {nx}x{nt} beta={beta}: plaquette {plaq:.5f} ({time.time() - t0:.1f}s)")
    n0 = nx * nt * 2
    B = np.stack([amgs.gaussian(n0, 1000 + i) for i in range(a.rhs)])
    dev = torch.device("cuda:0")
    Bd = torch.from_numpy(B.view(np.float64).copy()).to(dev)
    Xd = torch.zeros_like(Bd)

    def timed_solve(solver, nrhs, steps, warmup):
        its = np.zeros(nrhs, np.int32)
        rr = np.zeros(nrhs)
        ip = its.ctypes.data_as(amgs._I)
        rp = rr.ctypes.data_as(amgs._P)
        for _ in range(warmup):
            solver._chk(amgs.lib().amgs_solve_dev(solver.h, nrhs, Bd.data_ptr(), Xd.data_ptr(), ip, rp), "solve",
                        ok=(0, 1))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            solver._chk(amgs.lib().amgs_solve_dev(solver.h, nrhs, Bd.data_ptr(), Xd.data_ptr(), ip, rp), "solve",
                        ok=(0, 1))
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps, its.copy(), rr.copy()

    out = {}
    # ---- plain FP64, one right-hand side --------------------------------
    if not a.no_plain:
        sp = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, max_rhs=1), U)
        t = time.perf_counter()
        sp.setup()
        setup_plain = (time.perf_counter() - t) * 1e3
        ms_plain, it_plain, rr_plain = timed_solve(sp, 1, max(1, a.steps), a.warmup)
        log(f"plain FP64: setup {setup_plain:.0f} ms, solve {ms_plain:.1f} ms, iterations {it_plain[0]}, "
            f"relres {rr_plain[0]:.2e}")
        fp64_ms = sp.time_kernel(0, 1, 50)
        out.update(plain_fp64_ms_per_rhs=ms_plain, plain_setup_ms=setup_plain, plain_iterations=int(it_plain[0]),
                   plain_relres=float(rr_plain[0]))
        xs_plain = Xd[0].cpu().numpy().view(np.complex128).copy()
        sp.close()
    # ---- mixed precision, batched ----------------------------------------
    sm = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, max_rhs=a.rhs, precision=amgs.PREC_MIXED), U)
    t = time.perf_counter()
    sm.setup()
    setup_mixed = (time.perf_counter() - t) * 1e3
    ms_step, its, rrs = timed_solve(sm, a.rhs, a.steps, a.warmup)
    ms_rhs = ms_step / a.rhs
    log(f"mixed batched: setup {setup_mixed:.0f} ms, solve {ms_step:.1f} ms for {a.rhs} rhs -> {ms_rhs:.2f} ms/rhs, "
        f"iterations {its.min()}..{its.max()}, max relres {rrs.max():.2e}")
    if not a.no_plain:
        x0 = Xd[0].cpu().numpy().view(np.complex128)
        out["mixed_vs_fp64_x_relerr"] = float(np.linalg.norm(x0 - xs_plain) / np.linalg.norm(xs_plain))
        out["speedup_vs_plain_fp64"] = out["plain_fp64_ms_per_rhs"] / ms_rhs

    # ---- roofline: fine Wilson stencil (FP64, one rhs), HIP events on the library stream
    ns = nx * nt
    stencil_ms = sm.time_kernel(0, 1, 100)
    bytes_fp64 = 96.0 * ns  # psi 32 + U 32 + out 32 bytes per site
    achieved = bytes_fp64 / (stencil_ms * 1e-3) / 1e9
    traffic = None
    tr_path = os.path.join(ROOT, "profiles", "wilson_fp64_512_traffic.json")
    if os.path.exists(tr_path):
        traffic = json.load(open(tr_path)).get("bytes_per_launch")
    out["roofline"] = {"kernel": "k_wilson<double> (fine D, 512x512, 1 rhs)", "bound": "hbm",
                       "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                       "avg_launch_us": round(stencil_ms * 1e3, 3), "algorithmic_bytes_per_launch": bytes_fp64}

    # ---- CPU baseline: the oracle on this host, bounded sample -------------
    if not a.no_cpu:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        po = O.default_params(nx, nt, rows, m0=m0, threads=threads, out_m=a.cpu_iters, out_restarts=1)
        o = O.Oracle(po, U)
        t = time.perf_counter()
        o.setup(draw_order=1)
        t_setup = time.perf_counter() - t
        t = time.perf_counter()
        _, it_cpu, _, _ = o.solve(B[0])
        t_iters = time.perf_counter() - t
        per_iter = t_iters / max(it_cpu, 1)
        n_it = out.get("plain_iterations", int(its[0]))
        est = per_iter * n_it * 1e3
        log(f"oracle: setup {t_setup:.1f}s, {it_cpu} outer iterations in {t_iters:.1f}s -> {est / 1e3:.1f}s/rhs")
        out["cpu_baseline"] = {"value": round(est, 1), "unit": "ms/RHS", "cores": threads, "kind": "port",
                               "sample": f"oracle (C restatement, OpenMP) on {nx}x{nt}: set-up ({t_setup:.1f}s, "
                                         f"excluded) + {it_cpu} outer FGMRES iterations ({t_iters:.1f}s), scaled to "
                                         f"the {n_it} iterations of the FP64 solve"}

    res = {"metric": METRIC, "value": round(ms_rhs, 3), "unit": "ms/RHS", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(ms_step, 2), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32 inner / fp64 outer", "data": "synthetic",
           "config": {"workload": f"{nx}x{nt} 3-level K-cycle FGMRES(20) to 1e-10, beta={beta} Metropolis "
                                  f"({a.sweeps} sweeps), m0={m0}, rows 0 8 8 10 4 4 / 1 4 4 10 4 4",
                      "nrhs_batch": a.rhs, "levels": 3, "parallelism": "single GPU"},
           "iterations_min": int(its.min()), "iterations_max": int(its.max()), "relres_max": float(rrs.max()),
           "mixed_setup_ms": setup_mixed}
    res.update(out)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
