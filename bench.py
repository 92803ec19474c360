"""SIMPLE outer iterations per second on one MI355X (conjugate-heat-transfer
channel, 128x64x64 hexahedra x 6 jittered tetrahedra = 3,145,728 cells,
k-epsilon at Re=1e4, fluid energy + solid conduction, FP64).

A "step" is one SIMPLE outer iteration of the device-resident engine:
momentum (u, v, w) assembly + BiCGSTAB/Jacobi solves, Rhie-Chow, pressure
assembly + CG/ILU(0) solve, corrections, k and epsilon, fluid energy and the
coupled solid conduction solve. Inputs (mesh, geometry, fields) are resident
in HBM before the timed region; nothing leaves the device inside it except
the per-iteration residual scalars.

Prints ONE JSON line (rank 0). Extra keys: roofline (pressure CG SpMV
kernel, live HIP-event timing on the solver's stream), cpu_baseline (the
oracle, single core, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=1,
                    help="oracle outer iterations timed for cpu_baseline (0: skip)")
    ap.add_argument("--cpu-nx", type=int, default=0,
                    help="oracle sample mesh (0: same mesh as the GPU run)")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is a design, see the sharding section of DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from sysu_cfd_heat_amd import NSSolver, cases

    t0 = time.time()
    raw, p = cases.channel(nx=a.nx, ny=a.ny, nz=a.nz, tets=True, jitter=0.1, seed=1,
                           max_step=a.warmup + a.steps + 2)
    gpu = NSSolver(raw, p)
    setup_s = time.time() - t0
    print(f"[bench] setup {setup_s:.1f}s cells={raw.counts[1]} fluid={gpu.counts(0)[1]} "
          f"solid={gpu.counts(1)[1]}", file=sys.stderr, flush=True)
    for k in range(a.warmup):
        gpu.call("Iterate")
        print(f"[bench] warmup {k + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    gpu._fn("reset_stats")(gpu.h)
    gpu.profile(True)
    t1 = time.perf_counter()
    for k in range(a.steps):
        gpu.call("Iterate")  # returns after the device work of the iteration is done
        print(f"[bench] step {k + 1}/{a.steps} res={gpu.get('Residual')[:4].max():.3e}",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t1
    prof = gpu.profile_read()
    gpu.profile(False)
    st = gpu.stats()
    ms_per_step = 1000.0 * elapsed / a.steps
    value = a.steps / elapsed
    launches, tot_ms, bytes_per = prof[0], prof[1], prof[2]
    avg_ms = tot_ms / launches if launches else float("nan")
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9 if launches else None
    roof = {"bound": "hbm", "kernel": "k_cg_pq (pressure CG: p update + SpMV + p.q)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
            "launches": int(launches), "avg_us": avg_ms * 1e3,
            "bytes_per_launch": bytes_per}
    krylov = {"mom_iters_per_step": st[2] / a.steps, "p_iters_per_step": st[3] / a.steps,
              "ke_iters_per_step": st[4] / a.steps, "T_iters_per_step": st[5] / a.steps,
              "solid_iters_per_step": st[6] / a.steps}
    gpu.close()

    cpu = None
    if a.cpu_sample > 0:
        from oracle import oracle as O
        nx = a.cpu_nx or a.nx
        ny = a.ny if not a.cpu_nx else max(1, a.ny * a.cpu_nx // a.nx)
        nz = a.nz if not a.cpu_nx else max(1, a.nz * a.cpu_nx // a.nx)
        craw, cp = cases.channel(nx=nx, ny=ny, nz=nz, tets=True, jitter=0.1, seed=1,
                                 max_step=a.cpu_sample + 2)
        t2 = time.time()
        ref = O.NSSolver(craw, cp, mc=1)
        print(f"[bench] oracle setup {time.time() - t2:.1f}s", file=sys.stderr, flush=True)
        secs, _ = ref.time_iterations(a.cpu_sample)
        ref.close()
        cpu = {"value": a.cpu_sample / secs, "unit": "SIMPLE iters/s", "cores": 1,
               "kind": "port",
               "sample": f"first {a.cpu_sample} outer iteration(s) of the same case "
                         f"({nx}x{ny}x{nz}x6 tets), single thread, oracle",
               "seconds": secs}
        print(f"[bench] oracle {secs:.1f}s for {a.cpu_sample} iteration(s)", file=sys.stderr,
              flush=True)

    line = {
        "metric": "SIMPLE iters/s at 3.1M tets + SpMV HBM GB/s; p-solve iters learned vs plain",
        "value": value, "unit": "SIMPLE iters/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"CHT channel {a.nx}x{a.ny}x{a.nz} hex x6 jittered tets "
                               f"({raw.counts[1]} cells), k-eps Re=1e4, energy + solid",
                   "cells": raw.counts[1], "solvers": "BiCGSTAB+Jacobi / CG+ILU(0), 1e-8",
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "krylov": krylov,
        "device_s_per_step": st[1] / max(st[0], 1),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
