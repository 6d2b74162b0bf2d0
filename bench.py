"""Benchmark: SSP-RK3 steps of the Taylor-Green vortex on the wavy periodic 256^3 grid.

Workload: 257^3 nodes (256^3 unique: periodic faces are coincident end nodes), Mach 0.1, Re 1600,
Sutherland viscosity, Pr 0.72, characteristic WENO7 + AUSM+, 4th-order collocated viscous fluxes,
CFL 0.5 step computed on the device each step, FP64.  Each step is replayed from one hipGraph.
Prints one JSON line.

Further objects ride on the same line (skip with --no-reactive / --no-surrogate / --no-mixing):
  reactive   - the reacting H2/air periodic box (128^3 unique cells, 9 species, WENO5 + Roe, Strang-split
               batched exact chemistry with the autotuned launch shape), Mcell-steps/s;
  surrogate  - the chemistry surrogate trained on the GPU on exact-integrator states, then one chemistry
               sub-step of 256^3 cells by surrogate (with fallback) and by the exact integrator: speedup,
               max |dY|, max |dT| against the exact FP64 result, fallback fraction;
  mixing_layer - the reacting H2-N2/air mixing layer on 8 blocks of 384x192x96 in one context (WENO5 + HLLC,
               viscous, Strang chemistry, inflow noise, sponge outflow): Mcell-steps/s with exact chemistry and
               with the surrogate + fallback in the sub-steps, and one sub-step of a block both ways.
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
# flux kernel, per interface: FP64 operations (FMA counted twice) of WENO7-characteristic + AUSM+, as counted
# by SQ_INSTS_VALU_{FMA,MUL,ADD}_F64 on the 256^3 workload (42.3 GFLOP per launch over 257 x 257 x 258
# interfaces; profiles/flow_tgv256_fp64_valu.md)
FLUX_FLOP = 2480.0
FP64_PEAK_TFLOPS = 78.6
SCHEME_NAMES = {1: "WENO7-ch + LF", 2: "WENO7-ch + Roe", 3: "WENO7-ch + AUSM+", 4: "WENO7-ch + HLLC",
                51: "WENO5-cp LLF flux splitting", 52: "WENO5-ch LLF flux splitting",
                71: "WENO7-cp LLF flux splitting", 72: "WENO7-ch LLF flux splitting"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """nproc, the CPU model name and its clock, as the host reports them."""
    model, mhz = None, None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model is None:
                model = line.split(":", 1)[1].strip()
            elif line.startswith("cpu MHz") and mhz is None:
                mhz = float(line.split(":", 1)[1])
    except OSError:
        pass
    max_mhz = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("CPU max MHz"):
                max_mhz = float(line.split(":", 1)[1])
            elif line.startswith("Model name") and model is None:
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model,
            "mhz_now": mhz, "mhz_max": max_mhz}


def cpu_baseline(seconds=10.0, seconds_1core=6.0, n=65, cells_full=256 ** 3, steps_full=100):
    """The oracle (OpenMP C restatement of the reference's scheme routines) on a bounded sample of the same
    workload: the same scheme, grid family and initial field at n^3 nodes, timed with all host threads and
    with one.  Mcell-steps/s is a per-cell rate; the full workload's time at that rate is stated beside it."""
    import cases
    import oracle_lib
    from parity import make_pair
    case = cases.tgv(n, reynolds=1600.0, recon=5, inviscid=3, visc=4, cfl=0.5)
    # the oracle's state comes from its own initial condition path: reuse the product's initial bv
    S, B = make_pair(case)
    S.close()
    unique = (n - 1) ** 3
    lib = oracle_lib.load()
    threads = lib.orc_num_threads()

    def timed(limit, max_steps):
        B.step(B.cfl_dt())  # warm-up
        steps, t0 = 0, time.perf_counter()
        while True:
            B.step(B.cfl_dt())
            steps += 1
            el = time.perf_counter() - t0
            if el >= limit or steps >= max_steps:
                return steps, el

    steps, el = timed(seconds, 50)
    rate = unique * steps / el / 1e6
    lib.orc_set_threads(1)
    s1, e1 = timed(seconds_1core, 10)
    lib.orc_set_threads(threads)
    rate1 = unique * s1 / e1 / 1e6
    return {"value": round(rate, 4), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference scheme routines, OpenMP) on the same scheme and "
                      f"field at {n}^3 nodes ({n - 1}^3 unique), {steps} steps in {el:.1f} s on {threads} threads; "
                      f"one thread: {s1} steps in {e1:.1f} s",
            "value_1core": round(rate1, 5),
            "full_workload_s": round(cells_full * steps_full / (rate * 1e6), 1),
            "full_workload": f"{cells_full} cells x {steps_full} steps at this per-cell rate",
            "host": host_info()}


def reactive_bench(n=129, steps=50, warmup=5):
    """Strang-split reacting box: n^3 nodes ((n-1)^3 unique cells)."""
    import cases
    import corefl
    log(f"bench: reactive box {n}^3 nodes")
    case = cases.reactive_box(n, recon=4, inviscid=2, temporal=5, reaction=1, cfl=0.5)
    S = corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"])
    ngg = S.get_int("ngg")
    # 16 hot kernels, radius 4-10 cells (as the reacting-box configuration), velocity 20 m/s rms
    bv, sv = cases.reactive_state(n, ngg, seed=3, n_kernels=16, radius=(4.0, 10.0), k_max=8)
    S.set_state(bv, sv=sv)
    del bv, sv
    S.advance(warmup, graph=True)
    S.sync()
    # launch shape tuned on the state after the warm-up steps (the hot kernels are igniting: the stiff cells'
    # share is that of the timed steps); the step is captured again, untimed
    dt = S.compute_dt()
    ch = S.chemistry()
    ms = ch.autotune(0.5 * dt, reps=1)
    best = min(ms, key=ms.get)
    log(f"bench: chemistry autotune best {best} {ms[best]:.2f} ms (of {len(ms)} shapes)")
    S.advance(1, graph=True)
    S.sync()
    t0 = time.perf_counter()
    S.advance(steps, graph=True)
    S.sync()
    el = time.perf_counter() - t0
    unique = (n - 1) ** 3
    bvo = S.get(corefl.BV)
    T = bvo[5 * bvo.size // 6:]
    out = {"workload": f"H2/air periodic box {n - 1}^3 unique cells, 9 species / 19 reactions, WENO5 + Roe, "
                       "Strang-split batched ROK4E, 16 hot kernels, hipGraph replay",
           "value": round(unique * steps / el / 1e6, 3), "unit": "Mcell-steps/s", "steps": steps,
           "ms_per_step": round(el / steps * 1e3, 3), "chem_autotune": {"best": list(best), "ms": round(ms[best], 3)},
           "T_max": round(float(np.max(T)), 1), "finite": bool(np.isfinite(bvo).all())}
    log("bench: reactive " + json.dumps(out))
    S.close()
    return out


def surrogate_bench(n_states=4000000, train_steps=30000, n_eval=256, dt=5e-8, act="gelu",
                    gates=((0.5, 1e30), (0.5, 20.0), (0.5, 5.0), (0.5, 2.0), (0.35, 3.0), (0.35, 2.0), (0.25, 3.0),
                           (0.25, 2.0), (1.0, 10.0))):
    """Train the surrogate (12 -> 4 x 128 -> 10; act "gelu" as specified, or "hswish") on exact-integrator states,
    then time one chemistry sub-step of n_eval^3 cells by surrogate (with fallback, for each temperature gate) and
    by the exact integrator (autotuned launch)."""
    import chem_cases as CC
    import corefl
    import surrogate_train as ST
    t0 = time.perf_counter()
    ch = corefl.Chemistry(CC.MECH, CC.THERM)
    mw = ch.table("mw")
    data = ST.build_dataset(ch, n_states, 4, CC.Y_STOICH)
    t_data = time.perf_counter() - t0
    log(f"bench: surrogate dataset {data['T'].size} states in {t_data:.1f} s")
    sur = corefl.Surrogate(8192)
    sur.set_activation(corefl.Surrogate.ACT_GELU if act == "gelu" else corefl.Surrogate.ACT_HARD_SWISH)
    t0 = time.perf_counter()
    sc, loss = ST.train(sur, data, mw, steps=train_steps)
    t_train = time.perf_counter() - t0
    log(f"bench: surrogate trained in {t_train:.1f} s, loss {loss[0]:.3e} -> {np.mean(loss[-100:]):.3e}")
    del data
    rho, T, Y = CC.hot_kernel_box(n_eval, seed=3, n_kernels=16)
    ch.set_cells(rho, T, rho * Y)
    for _ in range(10):  # 20 us of exact chemistry: the hot kernels are igniting
        ch.integrate(2e-6)
    g = ch.get_cells()
    T1, Y1 = g["T"], g["rhoY"] / rho
    del g
    ch.set_cells(rho, T1, rho * Y1)
    ch.autotune(dt, reps=1)

    def timed(fn, reps=5):
        best, res = 1e30, None
        for _ in range(reps):
            ch.set_cells(rho, T1, rho * Y1)
            ch.sync()
            t0 = time.perf_counter()
            r = fn()
            ch.sync()
            best = min(best, (time.perf_counter() - t0) * 1e3)
            res = r
        return best, res

    exact_ms, _ = timed(lambda: ch.integrate(dt))
    x = ch.get_cells()
    sweep = []
    for gate, step in gates:
        ST.set_gate(sur, dT_max=gate, dT_step=step)
        ms, nf = timed(lambda: sur.chem_step(ch, dt))
        s = ch.get_cells()
        dy = float(np.max(np.abs(s["rhoY"] - x["rhoY"]) / rho))
        dT = float(np.max(np.abs(s["T"] - x["T"])))
        sweep.append({"gate_dT_K": gate, "gate_dT_step_K": step, "ms": round(ms, 3), "speedup": round(exact_ms / ms, 3), "max_abs_dY": dy,
                      "max_abs_dT_K": dT, "fallback_frac": round(nf / T1.size, 6)})
        log("bench: surrogate gate " + json.dumps(sweep[-1]))
    ok = [r for r in sweep if r["max_abs_dY"] <= 1e-3 and r["max_abs_dT_K"] <= 1.0]
    pick = max(ok, key=lambda r: r["speedup"]) if ok else min(sweep, key=lambda r: r["max_abs_dT_K"])
    out = {"workload": f"{n_eval}^3 cells of the hot-kernel H2/air box after 20 us of ignition, one chemistry "
                       f"sub-step of {dt:g} s",
           "network": "12 -> 4 x 128 -> 10, " + ("GELU (tanh form)" if act == "gelu" else "hard swish") +
                      ", BF16 weights, FP32 accumulation",
           "speedup": pick["speedup"], "gate_dT_K": pick["gate_dT_K"], "gate_dT_step_K": pick["gate_dT_step_K"],
           "accuracy_ok": bool(ok), "meets_targets": bool(ok) and pick["speedup"] >= 5.0,
           "max_abs_dY": pick["max_abs_dY"], "max_abs_dT_K": pick["max_abs_dT_K"],
           "fallback_frac": pick["fallback_frac"], "exact_ms": round(exact_ms, 3), "surrogate_ms": pick["ms"],
           "gate_sweep": sweep, "train_states": int(n_states), "train_steps": train_steps,
           "train_s": round(t_train, 1), "dataset_s": round(t_data, 1), "exact_launch": ch.get_launch(),
           "final_loss": float(np.mean(loss[-100:]))}
    return out


def stream_machs(ch, streams, names):
    """Mach numbers of the mixing layer's streams (speed U over the stream's sound speed, gamma from the mixture
    cp of the NASA-7 tables the chemistry context parsed)."""
    mw, lo, hi, tmid = ch.table("mw"), ch.table("lo"), ch.table("hi"), ch.table("t_mid")
    out = []
    for st in streams:
        Y = np.array([st["Y"].get(s, 0.0) for s in names])
        T = st["T"]
        a = np.where((T < tmid)[:, None], lo, hi)
        cp_r = a[:, 0] + a[:, 1] * T + a[:, 2] * T ** 2 + a[:, 3] * T ** 3 + a[:, 4] * T ** 4
        R = 8314.462618 * np.sum(Y / mw)
        cp = float(np.sum(Y * cp_r * 8314.462618 / mw))
        out.append(st["U"] / np.sqrt(cp / (cp - R) * R * T))
    return out


def mixing_layer_bench(n_blocks=8, shape=(384, 192, 96), steps=3, warmup=1, n_states=1000000, train_steps=15000):
    """The reacting H2-N2 / air mixing layer on n_blocks blocks joined in a row (WENO5 + HLLC, viscous, Strang-split
    chemistry, sponge outflow, seeded inflow noise), SSP-RK3 steps replayed from a hipGraph: Mcell-steps/s with the
    batched exact chemistry and with the surrogate (gate + exact fallback) in the sub-steps, and one chemistry
    sub-step of a block's cells both ways against the exact FP64 result."""
    import cases
    import chem_cases as CC
    import corefl
    import surrogate_train as ST
    names = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    ch0 = corefl.Chemistry(CC.MECH, CC.THERM)
    machs = stream_machs(ch0, (cases.MIXING_UPPER, cases.MIXING_LOWER), names)
    mx, my, mz = shape
    t0 = time.perf_counter()
    case = cases.mixing_layer_blocks(n_blocks=n_blocks, mx=mx, my=my, mz=mz, mach=machs, recon=4, inviscid=4, visc=2,
                                     reaction=1, temporal=5)
    S = corefl.Solver(setup_text=case["setup"], blocks=case["grids"], boundary_text=case["boundary"],
                      inner_text=case["inner"])
    del case["grids"]
    cells = case["cells"]
    log(f"bench: mixing layer {n_blocks} x {mx}x{my}x{mz} ready in {time.perf_counter() - t0:.1f} s")
    S.advance(warmup, graph=True)
    S.sync()

    def timed_steps():
        S.advance(1, graph=True)  # (re)capture, untimed
        S.sync()
        t1 = time.perf_counter()
        S.advance(steps, graph=True)
        S.sync()
        return time.perf_counter() - t1

    el_exact = timed_steps()
    log(f"bench: mixing layer exact chemistry {el_exact / steps * 1e3:.1f} ms/step")
    # the surrogate, trained on the exact integrator's states of the two streams' mixtures
    t1 = time.perf_counter()
    ya = np.array([cases.MIXING_UPPER["Y"].get(s, 0.0) for s in names])
    yb = np.array([cases.MIXING_LOWER["Y"].get(s, 0.0) for s in names])
    data = ST.build_mixing_dataset(ch0, n_states, 5, ya, yb, cases.MIXING_UPPER["T"], cases.MIXING_LOWER["T"])
    sur = corefl.Surrogate(8192)
    _, loss = ST.train(sur, data, ch0.table("mw"), steps=train_steps)
    del data
    ST.set_gate(sur, dT_max=0.35, dT_step=3.0)
    t_train = time.perf_counter() - t1
    log(f"bench: mixing-layer surrogate trained in {t_train:.1f} s, loss {np.mean(loss[-100:]):.3e}")
    # one chemistry sub-step of the middle block's cells: surrogate with fallback against the exact sub-step
    dt = S.compute_dt()
    ch = S.chemistry(n_blocks // 2)
    g = ch.get_cells()
    rho = ch.get_cells()["rhoY"].sum(axis=0)
    T0, Y0 = g["T"], g["rhoY"] / rho
    del g

    def best_of(fn, reps=3):
        best = 1e30
        for _ in range(reps):
            ch.set_cells(rho, T0, rho * Y0)
            ch.sync()
            t2 = time.perf_counter()
            fn()
            ch.sync()
            best = min(best, time.perf_counter() - t2)
        return best * 1e3

    ms_exact = best_of(lambda: ch.integrate(0.5 * dt))
    x = ch.get_cells()
    nf = [0]
    ms_sur = best_of(lambda: nf.__setitem__(0, sur.chem_step(ch, 0.5 * dt)))
    sg = ch.get_cells()
    sub = {"cells": int(T0.size), "dt_s": 0.5 * dt, "exact_ms": round(ms_exact, 3), "surrogate_ms": round(ms_sur, 3),
           "speedup": round(ms_exact / ms_sur, 3),
           "max_abs_dY": float(np.max(np.abs(sg["rhoY"] - x["rhoY"]) / rho)),
           "max_abs_dT_K": float(np.max(np.abs(sg["T"] - x["T"]))), "fallback_frac": round(nf[0] / T0.size, 6)}
    del x, sg
    log("bench: mixing-layer chemistry sub-step " + json.dumps(sub))
    S.set_chemistry_surrogate(sur)
    el_sur = timed_steps()
    bv = S.get(corefl.BV, n_blocks // 2)
    finite = bool(np.isfinite(bv).all())
    T = bv[5 * bv.size // 6:]
    out = {"workload": f"reacting H2-N2/air mixing layer, {n_blocks} blocks of {mx}x{my}x{mz} joined in a row "
                       f"({cells} nodes), 9 species / 19 reactions, WENO5 + HLLC, 2nd-order viscous with mixture "
                       "transport, Strang-split chemistry, inflow noise 2%, sponge outflow, hipGraph replay",
           "cells": cells, "steps": steps,
           "exact": {"value": round(cells * steps / el_exact / 1e6, 3), "unit": "Mcell-steps/s",
                     "ms_per_step": round(el_exact / steps * 1e3, 2)},
           "surrogate": {"value": round(cells * steps / el_sur / 1e6, 3), "unit": "Mcell-steps/s",
                         "ms_per_step": round(el_sur / steps * 1e3, 2)},
           "step_speedup": round(el_exact / el_sur, 3), "chem_substep": sub, "train_states": n_states,
           "train_steps": train_steps, "train_s": round(t_train, 1), "T_max_mid_block": round(float(np.max(T)), 1),
           "finite": finite}
    log("bench: mixing layer " + json.dumps(out))
    S.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=257, help="nodes per direction (n-1 unique)")
    ap.add_argument("--inviscid", type=int, default=3,
                    help="inviscid_scheme of the Taylor-Green workload: 3 (AUSM+ after characteristic WENO7, the "
                         "default) or 51/52/71/72 (LLF flux-splitting WENO5/WENO7, component-wise/characteristic)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=3)
    ap.add_argument("--no-reactive", action="store_true")
    ap.add_argument("--reactive-steps", type=int, default=50, help="timed steps of the reacting box (50 as specified)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-mixing", action="store_true", help="skip the multiblock reacting mixing layer")
    ap.add_argument("--mixing-blocks", type=int, default=8)
    ap.add_argument("--mixing-shape", type=int, nargs=3, default=(384, 192, 96))
    ap.add_argument("--mixing-steps", type=int, default=3)
    ap.add_argument("--only", choices=("reactive", "surrogate", "mixing"),
                    help="run one secondary section alone and print its JSON object (profiling runs)")
    ap.add_argument("--train-steps", type=int, default=30000)
    ap.add_argument("--surrogate-states", type=int, default=4000000)
    ap.add_argument("--surrogate-act", choices=("gelu", "hswish"), default="gelu",
                    help="hidden activation of the surrogate (GELU is the specified network)")
    ap.add_argument("--surrogate-hardswish", action="store_true",
                    help="also train and time the hard-swish network, reported as an extra object")
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

    if args.only == "reactive":
        print(json.dumps(reactive_bench(steps=args.reactive_steps)), flush=True)
        return
    if args.only == "mixing":
        print(json.dumps(mixing_layer_bench(args.mixing_blocks, tuple(args.mixing_shape), args.mixing_steps)),
              flush=True)
        return
    if args.only == "surrogate":
        print(json.dumps(surrogate_bench(n_states=args.surrogate_states, train_steps=args.train_steps,
                                         act=args.surrogate_act)), flush=True)
        return
    n = args.n
    log(f"bench: building TGV wavy grid {n}^3")
    inv = args.inviscid
    case = cases.tgv(n, reynolds=1600.0, recon=5, inviscid=inv, visc=4, cfl=0.5)
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
    d = "xyz".index(dom[-1])
    symbol = (f"k_inviscid<{d}, 4, {inv}>" if inv < 50 else
              f"k_fsweno<{d}, {4 if inv > 70 else 3}, {'true' if inv % 10 == 2 else 'false'}>")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom, "symbol": symbol,
                "algorithmic_bytes": FLUX_BYTES * faces}
    # the flux kernels are bound by FP64 arithmetic, not by HBM: their FP64 VALU rate against the peak
    fp64 = None
    if inv == 3:
        tflops = FLUX_FLOP * faces / (flux[dom] * 1e-3) / 1e12
        fp64 = {"kernel": dom, "achieved": round(tflops, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / FP64_PEAK_TFLOPS, 4), "flop_per_interface": FLUX_FLOP}

    S.close()
    reactive = None if args.no_reactive else reactive_bench(steps=args.reactive_steps)
    surrogate = None if args.no_surrogate else surrogate_bench(n_states=args.surrogate_states,
                                                               train_steps=args.train_steps, act=args.surrogate_act)
    surrogate_hs = (surrogate_bench(n_states=args.surrogate_states, train_steps=args.train_steps, act="hswish")
                    if args.surrogate_hardswish and not args.no_surrogate else None)
    mixing = None if args.no_mixing else mixing_layer_bench(args.mixing_blocks, tuple(args.mixing_shape),
                                                             args.mixing_steps)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle)")
        cpu = cpu_baseline()
    step_frac = value * 1e6 * STEP_BYTES / (HBM_PEAK_GBS * 1e9)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"taylor-green wavy-grid {n - 1}^3 unique cells ({n}^3 nodes), {SCHEME_NAMES[inv]}, "
                               "4th-order viscous, SSP-RK3, CFL 0.5 device step, hipGraph replay",
                   "cells": unique, "reynolds": 1600, "mach": 0.1},
        "step_hbm_frac": round(step_frac, 4),
        "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
        "roofline": roofline,
        "fp64_valu": fp64,
        "cpu_baseline": cpu,
        "reactive": reactive,
        "surrogate": surrogate,
        "mixing_layer": mixing,
    }
    if surrogate_hs is not None:
        line["surrogate_hard_swish"] = surrogate_hs  # an alternative network, not the specified one
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
