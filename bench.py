"""Throughput of the VNR Lagrange + remap step on one MI355X.

A step is one full iteration of the reference time loop (CFL reduction,
Lagrange phase, two remap sweeps with the dual remap, final remap + EOS)
over the whole mesh; the metric is cell-updates/s = cells x steps / wall
time of the timed steps, with the state resident in HBM.

Default workload: the point explosion on 4096^2 (one material, minmod
order 2) in the plain FP64 form, the bandwidth-roofline run.  The same line
carries, under "ensemble", the 512-problem ensemble (sides 64..1024, seed
20261204) stepped batched in the mixed FP32 / FP64 form against the same
problems stepped one after another in the plain FP64 form (--no-ensemble
skips it).  --config c2 runs the two-material Riemann problem on 2048^2
(superBee + PenteBorne); --config c4 runs the three-material layout in the
mixed form with a per-field error report against the FP64 form; --config c5
runs the ensemble alone; --config c1 runs the reference's own test input
(Sod X on 200x5), whose 1000 cells leave the step bound by launch latency;
--config euc runs the Eucclhyd scheme (the second Lagrange scheme over the
same cell sweeps) on the reference's two-material Sod input at 1024^2.

Prints one JSON line.  --gpus N > 1 exits: the library runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s (Lagrange+remap) + HBM GB/s fraction; mixed/batched vs plain FP64"
HBM_PEAK = 8.0e12


PRECISION_NAME = {0: "FP64", 1: "mixed FP32/FP64"}


def progress(msg):
    """One line on stderr at each phase boundary (stdout carries only the
    JSON result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
DTYPE = {0: "f64", 1: "f32+f64"}


def workload(name):
    from mahyco_amd import configs
    if name == "c1":
        import mahyco_amd
        path = os.path.join(ROOT, "tests", "golden", "vnr", "SOD_X_Mono_mat_Minmod_vnr_o2", "donnees.txt")
        return "Sod X 200x5, 1 material, minmod order 2 (the reference's own test input)", mahyco_amd.read_donnees(path)
    if name == "c2":
        return "two-material Riemann 2048x2048, perturbed interface, superBee+PenteBorne", configs.riemann_perturbed()
    if name == "c3":
        return "point explosion 4096x4096, 1 material, minmod order 2", configs.point_explosion()
    if name == "c4":
        return "three-material layout 3072x1024, minmod, density jitter 1e-3", configs.triple_layout()
    if name == "euc":
        # the second Lagrange scheme over the same cell sweeps: the reference's
        # two-material Sod input with PenteBorne-mixte, on 1024^2 (its fixed
        # first time step, 1e-4, is about the CFL step of this mesh)
        import mahyco_amd
        path = os.path.join(ROOT, "tests", "golden", "euc", "SOD_X_bi_mat_Superbee_PenteBorne_mix", "donnees.txt")
        cfg = mahyco_amd.read_donnees(path)
        cfg.nx = cfg.ny = 1024
        cfg.dx = cfg.dy = 1.0 / 1024
        return "Eucclhyd two-material Sod 1024x1024, superBee + PenteBorne-mixte", cfg
    if name.startswith("riemann@"):
        n = int(name[8:])
        c = configs.riemann(n, 0.45, 1.0, 1.0, 0.125, 0.1, 1.4, 1.67)
        return f"two-material Riemann {n}x{n}, straight interface, minmod", c
    if name.startswith("c2@"):
        n = int(name[3:])
        c = configs.riemann_perturbed(n=n)
        return f"two-material Riemann {n}x{n}, perturbed interface, superBee+PenteBorne", c
    if name.startswith("c4@"):
        nx = int(name[3:])
        return f"three-material layout {nx}x{nx // 3}, minmod, density jitter 1e-3", configs.triple_layout(nx, nx // 3)
    raise SystemExit(f"unknown config {name}")


def algorithmic_bytes_per_cell(m, precision=0):
    """Compulsory persistent state per cell-update, read once and written
    once: cell rho, p, q, c, divu + 7 words per material (f_k, mu_k, rho_k,
    e_k, p_k, q_k, M_k), mixture e written, node u_x, u_y, M, and the
    mixed/pure flags (4 B read + 4 B written each).  FP64: 8 B words; mixed
    form: f_k and M_k 8 B, the rest 4 B."""
    if precision == 1:
        return 4 * (17 + 10 * m) + 32 * m + 16
    return 8 * (17 + 14 * m) + 16


def cpu_baseline(cfg, seconds_budget):
    """The CPU oracle (same arithmetic, OpenMP over cells) on the same mesh
    for a bounded number of steps."""
    from oracle import refcpu
    cores = os.cpu_count() or 1
    refcpu.set_threads(cores)
    progress(f"cpu baseline: oracle set-up {cfg.nx}x{cfg.ny}, {cores} threads")
    orc = refcpu.Oracle(cfg)
    orc.step(1)  # warm-up step (first-touch of the arrays)
    progress("cpu baseline: timed steps")
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.step(1)
        n += 1
        if time.perf_counter() - t0 > seconds_budget or n >= 20:
            break
    dt = time.perf_counter() - t0
    orc.close()
    cells = cfg.nx * cfg.ny
    return {"value": cells * n / dt, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"CPU oracle (reference-semantics restatement, g++ -O3 -fopenmp), same {cfg.nx}x{cfg.ny} "
                      f"mesh, {n} timed steps after 1 warm-up step"}


def precision_report(cfg, steps):
    """Per-field relative L1 / Linf errors of the mixed form against the FP64
    form after `steps` steps from the same initial state, the two forms'
    cell-update rates over those steps, and the total-mass drift of each."""
    import mahyco_amd
    from mahyco_amd import precision
    rates, ctxs, mass0 = {}, {}, None
    for prec in (0, 1):
        cfg.precision = prec
        ctx = mahyco_amd.Context(cfg)
        if mass0 is None:
            mass0 = ctx.conservation()[0]
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run(steps)
        rates[prec] = cfg.nx * cfg.ny * steps / (time.perf_counter() - t0)
        ctxs[prec] = ctx
    errs = precision.field_errors(ctxs[1], ctxs[0], cfg.nbmat)
    drift = {DTYPE[p]: abs(ctxs[p].conservation()[0] - mass0) / mass0 for p in (0, 1)}
    t = ctxs[0].time_state().t_n
    for c in ctxs.values():
        c.close()
    cfg.precision = 1
    return {"steps": steps, "t_end": t, "against": "FP64 form (bitwise equal to the CPU oracle)",
            "fields": {k: {"rel_L1": v[0], "rel_Linf": v[1]} for k, v in errs.items()},
            "mass_drift": drift, "fp64_rate": rates[0], "mixed_rate": rates[1],
            "mixed_over_fp64": rates[1] / rates[0]}


def copy_probe(mib):
    """A plain device-to-device copy of known size (the runtime's copy
    kernel), used to calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n, "copies": 3}), flush=True)


def ensemble_rates(args):
    """C5: 512 two-material Riemann problems of uneven sizes stepped together
    (one launch per kernel over all problems, two captured HIP graphs), in
    the mixed form unless --precision fp64, against the same problems stepped
    one after another with the plain single-problem FP64 path.  Both legs
    replay captured HIP graphs in their timed steps: the batch captures its
    two graphs in its warm-up, and each plain context runs two untimed steps
    first, which capture its two graphs (one per step parity)."""
    import mahyco_amd
    from mahyco_amd import configs
    prec = 1 if args.precision in (None, "mixed") else 0
    cfgs = configs.ensemble(n=args.ensemble)
    for c in cfgs:
        c.precision = prec
    progress(f"ensemble: {len(cfgs)} problems, batched set-up")
    batch = mahyco_amd.Batch(cfgs)
    cells = batch.cells
    batch.run(max(args.warmup, 12))  # the first 10 steps also tune the launch widths
    progress("ensemble: batched timed steps")
    t0 = time.perf_counter()
    batch.run(args.steps)  # ends with a stream synchronisation
    wall = time.perf_counter() - t0
    batch.close()
    value = cells * args.steps / wall
    plain_steps = max(1, min(args.steps, args.plain_steps))
    plain_t = 0.0
    progress("ensemble: plain FP64 problems one after another")
    for n, c in enumerate(cfgs):
        if n and n % 128 == 0:
            progress(f"ensemble: plain problem {n}")
        c.precision = 0
        ctx = mahyco_amd.Context(c)
        ctx.run(2)  # captures the graphs of both step parities
        t1 = time.perf_counter()
        ctx.run(plain_steps)
        plain_t += time.perf_counter() - t1
        ctx.close()
    plain = cells * plain_steps / plain_t
    return {"workload": f"ensemble of {len(cfgs)} two-material Riemann problems, sides 64..1024 log-uniform "
                        f"(seed 20261204), minmod, batched, {PRECISION_NAME[prec]}",
            "problems": len(cfgs), "cells": cells, "dtype": DTYPE[prec], "steps": args.steps,
            "value": value, "unit": "cell-updates/s", "ms_per_step": wall / args.steps * 1e3,
            "plain_form": {"value": plain, "unit": "cell-updates/s", "steps": plain_steps,
                           "note": "same problems, FP64, one context after another (mhy_run per problem, "
                                   "graphs captured before the timed steps)"},
            "batched_over_plain": value / plain}


def bench_ensemble(args):
    """--config c5: the ensemble comparison as the line itself."""
    r = ensemble_rates(args)
    prec = 1 if r["dtype"] == DTYPE[1] else 0
    B = algorithmic_bytes_per_cell(2, prec)
    out = {
        "metric": METRIC, "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": r["dtype"], "data": "synthetic",
        "config": {"workload": r["workload"], "problems": r["problems"], "cells": r["cells"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "whole step (batched)", "achieved": B * r["value"] / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": B * r["value"] / HBM_PEAK, "traffic": None},
        "plain_form": r["plain_form"],
        "batched_over_plain": r["batched_over_plain"],
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"))
    ap.add_argument("--ensemble", type=int, default=512, help="problems of the c5 ensemble")
    ap.add_argument("--plain-steps", type=int, default=5, help="steps of the one-by-one comparison (c5)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="leave out the ensemble comparison the default (c3) line carries")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="fp64 or mixed FP32/FP64 (default: mixed for c4 and c5, fp64 otherwise)")
    ap.add_argument("--report-steps", type=int, default=500,
                    help="steps of the per-field error report against FP64 (mixed form)")
    ap.add_argument("--kernel-timing", choices=["inline", "after"], default="inline",
                    help="HIP events around the dominant kernel's launches in the timed steps (inline), or "
                         "around every launch of a few steps after them (after)")
    ap.add_argument("--projection-order", type=int, choices=[1, 2, 3], default=None,
                    help="override the workload's remap order (3: third-order faces, FP64 form)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the library runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import mahyco_amd

    if args.config == "c5":
        return bench_ensemble(args)
    desc, cfg = workload(args.config)
    if args.precision is None:
        args.precision = "mixed" if args.config.startswith("c4") else "fp64"
    prec = 1 if args.precision == "mixed" else 0
    cfg.precision = prec
    desc = f"{desc}, {PRECISION_NAME[prec]}"
    traffic_key = args.config
    if args.projection_order is not None and args.projection_order != cfg.projection_order:
        cfg.projection_order = args.projection_order
        desc = f"{desc}, remap order {args.projection_order}"
        traffic_key = f"{args.config}_o{args.projection_order}"
    progress(f"{desc}: set-up")
    ctx = mahyco_amd.Context(cfg)
    cells = cfg.nx * cfg.ny
    # the launch-shape tuner's steps (meshes of at least 2^20 cells) are part
    # of the set-up: they time the interchangeable y kernels before any
    # graph is captured, and come before the warm-up
    tuning_steps = 0
    if cells >= (1 << 20) and cfg.schema != 2000:  # (the Eucclhyd scheme has no tuner)
        while not ctx.tuned_kernels() and tuning_steps < 64:
            ctx.run(1)
            tuning_steps += 1
    progress("warm-up and timed steps")
    ctx.run(args.warmup)
    ctx.synchronize()
    if args.kernel_timing == "inline":
        # an untimed pass with HIP events around every launch names the
        # dominant kernel (and fills the event pool); the timed steps then
        # record events on the context stream around that kernel only
        ctx.enable_timing(True)
        ctx.run(args.steps)
        stats = ctx.kernel_stats()
        tot = {k: v["avg_ms"] * v["launches"] for k, v in stats.items()}
        dom = max(tot, key=tot.get)
        ctx.set_timing_mask(1 << stats[dom]["id"])
        ctx.enable_timing(True)  # clears the statistics, keeps the events
    t0 = time.perf_counter()
    ctx.run(args.steps)  # ends with a stream synchronisation
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall
    if args.kernel_timing == "inline":
        stats[dom] = ctx.kernel_stats()[dom]  # measured over the timed steps
    else:
        # the same events on a few steps after the timed ones
        ctx.enable_timing(True)
        ctx.run(max(3, min(args.steps, 10)))
        stats = ctx.kernel_stats()
        tot = {k: v["avg_ms"] * v["launches"] for k, v in stats.items()}
        dom = max(tot, key=tot.get)
    ctx.enable_timing(False)
    tuned = ctx.tuned_kernels()
    ctx.close()
    dk = stats[dom]
    achieved = dk["bytes"] / (dk["avg_ms"] * 1e-3) if dk["avg_ms"] > 0 else 0.0
    traffic = None
    if os.path.exists(args.traffic_json):
        try:
            with open(args.traffic_json) as fh:
                traffic = json.load(fh).get(traffic_key, {}).get(dom)
        except (OSError, ValueError):
            traffic = None
    B = algorithmic_bytes_per_cell(cfg.nbmat, prec)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[prec],
        "data": "synthetic",
        "config": {"workload": desc, "mesh": [cfg.nx, cfg.ny], "nbmat": cfg.nbmat, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic},
        # the compulsory-state count below is the VNR state's
        "step_roofline": None if cfg.schema == 2000 else {
            "algorithmic_bytes_per_cell_update": B, "achieved": B * value / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": B * value / HBM_PEAK},
        "kernels": {k: {"avg_ms": round(v["avg_ms"], 5), "GBps_alg": round(v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9, 1)
                        if v["avg_ms"] > 0 else 0} for k, v in stats.items() if v["launches"]},
        "tuned_kernels": tuned or None,
        "tuning_steps": tuning_steps,
        "kernel_timing": ("HIP events on the context stream: the dominant kernel over the timed steps, the others "
                          "over an untimed pass just before" if args.kernel_timing == "inline" else
                          "HIP events on the context stream over a few steps after the timed ones"),
    }
    if prec == 1 and args.report_steps > 0:
        out["precision_report"] = precision_report(cfg, args.report_steps)
    if args.config == "c1":
        # a mesh this small is bound by launch latency: the same problem as a
        # batch of one replays each step as a captured HIP graph
        batch = mahyco_amd.Batch([cfg])
        batch.run(max(args.warmup, 12))  # the first 10 steps tune the launch widths
        t0 = time.perf_counter()
        batch.run(args.steps)
        gwall = time.perf_counter() - t0
        batch.close()
        out["graph_form"] = {"value": cells * args.steps / gwall, "ms_per_step": gwall / args.steps * 1e3,
                             "note": "the same problem as a batch of one (captured HIP graphs)"}
    if not args.no_cpu_baseline:
        cfg.precision = 0
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_seconds)
    if args.config == "c3" and not args.no_ensemble:
        ens_args = argparse.Namespace(**vars(args))
        ens_args.precision = None  # the ensemble's own form: mixed, against plain FP64
        out["ensemble"] = ensemble_rates(ens_args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
