"""Benchmark: SSP-RK3 steps of the Taylor-Green vortex on the wavy periodic 256^3 grid.

Workload: 257^3 nodes (256^3 unique: periodic faces are coincident end nodes), Mach 0.1, Re 1600,
Sutherland viscosity, Pr 0.72, characteristic WENO7 + AUSM+, 4th-order collocated viscous fluxes,
CFL 0.5 step computed on the device each step, FP64.  Each step is replayed from one hipGraph.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "corefl-cpc_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Mcell-steps/s per RK3 step + %HBM roofline; surrogate vs exact chem speedup"
HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per cell-step of the nv = 5 path: 3 stages x (U in 40 B + metrics and J 80 B +
# U out 40 B) + 2 x U^n 40 B
STEP_BYTES = 560.0
# flux kernel, per interface: conservative state (40 B) + 3 metric components and J (32 B) read, 5 fluxes
# (40 B) written
FLUX_BYTES = 112.0


def measured_traffic(kernel_symbol, n):
    """HBM bytes per launch of one kernel from the committed rocprofv3 counter summary (same workload only)."""
    path = os.path.join(ROOT, "profiles", "flow_tgv256_hbm_traffic.md")
    if n != 257 or not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 7 and cells[1] == f"`{kernel_symbol}`":
            return float(cells[7]) * 1e9
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds=12.0, n=65):
    """The oracle (OpenMP C restatement) on a bounded sample of the same workload."""
    import cases
    import oracle_lib
    from parity import make_pair
    case = cases.tgv(n, reynolds=1600.0, recon=5, inviscid=3, visc=4, cfl=0.5)
    # the oracle's state comes from its own initial condition path: reuse the product's initial bv
    S, B = make_pair(case)
    S.close()
    unique = (n - 1) ** 3
    B.step(B.cfl_dt())  # warm-up
    steps, t0 = 0, time.perf_counter()
    while True:
        B.step(B.cfl_dt())
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    return {"value": unique * steps / el / 1e6, "unit": "Mcell-steps/s", "cores": oracle_lib.load().orc_num_threads(),
            "kind": "port",
            "sample": f"oracle (C, OpenMP) on the same scheme at {n}^3 nodes ({n - 1}^3 unique), {steps} steps, "
                      f"{el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=257, help="nodes per direction (n-1 unique)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reference point for the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is designed, not built.",
              file=sys.stderr)
        sys.exit(2)

    import cases
    import corefl

    n = args.n
    log(f"bench: building TGV wavy grid {n}^3")
    case = cases.tgv(n, reynolds=1600.0, recon=5, inviscid=3, visc=4, cfl=0.5)
    t0 = time.perf_counter()
    S = corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"])
    del case["grid"]
    log(f"bench: context ready in {time.perf_counter() - t0:.1f} s; warmup {args.warmup} steps")
    S.advance(args.warmup, graph=True)
    S.sync()
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    S.advance(args.steps, graph=True)
    S.sync()
    el = time.perf_counter() - t0
    unique = (n - 1) ** 3
    value = unique * args.steps / el / 1e6
    bv = S.get(corefl.BV)
    finite = bool(np.isfinite(bv).all())
    log(f"bench: {value:.1f} Mcell-steps/s, {el / args.steps * 1e3:.2f} ms/step, finite={finite}")
    if not finite:
        raise SystemExit("bench: non-finite state")

    ms = S.profile_kernels(args.profile_reps)
    log("bench: kernel ms " + json.dumps({k: round(v, 4) for k, v in ms.items()}))
    flux = {k: ms[k] for k in ("flux_x", "flux_y", "flux_z")}
    dom = max(flux, key=flux.get)
    faces = {"flux_x": (n + 1) * n * n, "flux_y": n * (n + 1) * n, "flux_z": n * n * (n + 1)}[dom]
    achieved = FLUX_BYTES * faces / (flux[dom] * 1e-3) / 1e9
    symbol = {"flux_x": "k_inviscid<0, 4, 3>", "flux_y": "k_inviscid<1, 4, 3>", "flux_z": "k_inviscid<2, 4, 3>"}[dom]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(symbol, n), "kernel": dom,
                "algorithmic_bytes": FLUX_BYTES * faces}

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle)")
        cpu = cpu_baseline()
    step_frac = value * 1e6 * STEP_BYTES / (HBM_PEAK_GBS * 1e9)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"taylor-green wavy-grid {n - 1}^3 unique cells ({n}^3 nodes), WENO7-ch + AUSM+, "
                               "4th-order viscous, SSP-RK3, CFL 0.5 device step, hipGraph replay",
                   "cells": unique, "reynolds": 1600, "mach": 0.1},
        "step_hbm_frac": round(step_frac, 4),
        "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
