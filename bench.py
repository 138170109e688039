"""Benchmark: FP64 cell-stage updates/s of the SSP3 rate-of-change hot path on
one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One step = one SSP3 step (3 stages) of the whole instance, replayed from a
captured HIP graph with the state resident in HBM.  Cell-stage updates/s =
n_cells * 3 * K / T.  Prints one JSON line (see README of the contract in the
task description): value, roofline of the dominant kernel (reconstruction,
HIP-event timed on the handle's stream) and the CPU baseline (the oracle, a
restatement of the reference's OpenMP CPU path, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def build_workload(name, args):
    from zisafvm_amd import workloads as W

    if name == "vortex2d":
        return W.vortex_2d(n=args.n or 128)
    if name == "isothermal2d":
        return W.isothermal_2d(n=args.n or 360)
    if name == "polytrope3d":
        return W.polytrope_3d(n=args.n or 40)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(wl, budget_s):
    """Oracle RoC calls on this host's cores for about budget_s seconds."""
    from oracle.zfv_ref import Oracle

    cores = cpu_cores()
    o = Oracle(wl.mesh, wl.params, threads=cores)
    if wl.steady is not None:
        o.set_steady_state(wl.steady)
    t0 = time.perf_counter()
    o.rate(wl.u0)
    first = time.perf_counter() - t0
    calls = max(1, int(budget_s / max(first, 1e-3)))
    calls = min(calls, 20)
    t0 = time.perf_counter()
    for _ in range(calls):
        o.rate(wl.u0)
    el = time.perf_counter() - t0
    n = wl.mesh.n_cells
    return {
        "value": n * calls / el,
        "unit": "cell-stage updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{calls} rate-of-change evaluations (one RK stage each) of the same {wl.name} instance "
        f"({n} cells), oracle with OpenMP on {cores} threads, {el:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="isothermal2d")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run only the streaming-copy calibration kernel (1 GiB x 5) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is a design "
              f"in DESIGN.md, not part of this build.", file=sys.stderr)
        return 2

    import zisafvm_amd as z

    if args.copy_probe_only:
        bw = z.bandwidth_probe(1 << 30, 5)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "iters": 5, "gbytes_per_s": bw}), flush=True)
        return 0

    t0 = time.perf_counter()
    wl = build_workload(args.workload, args)
    t_mesh = time.perf_counter() - t0
    log(f"workload {wl.name}: {wl.mesh.n_cells} cells, mesh+IC {t_mesh:.1f}s")
    t0 = time.perf_counter()
    s = z.Solver(wl.mesh, wl.params)
    t_setup = time.perf_counter() - t0
    s.set_state(wl.u0)
    if wl.steady is not None:
        s.set_steady_state(wl.steady)
    info = s.info()
    log(f"setup {t_setup:.1f}s, faces {info.n_faces}, device bytes {info.device_bytes/1e9:.2f} GB")
    dt = s.cfl_dt(wl.cfl)
    s.steps(args.warmup, 0.0, dt)
    s.synchronize()
    t0 = time.perf_counter()
    s.steps(args.steps, 0.0, dt)
    s.synchronize()
    el = time.perf_counter() - t0
    n = wl.mesh.n_cells
    cell_stages = n * 3 * args.steps
    value = cell_stages / el
    ms_per_step = 1e3 * el / args.steps
    log(f"{args.steps} steps in {el:.3f}s: {value:.3e} cell-stage updates/s")

    # per-kernel HIP-event timing on the handle's stream
    kt = s.time_kernels(max(2, min(10, args.steps)), dt)
    launches = kt.n_launches
    rec_ms = kt.reconstruct_ms / launches
    face_ms = kt.face_flux_ms / launches
    upd_ms = kt.update_ms / launches
    rec_bytes = info.recon_bytes_per_cell * n
    achieved = rec_bytes / (rec_ms * 1e-3) / 1e9
    log(f"kernel avg ms: reconstruct {rec_ms:.4f} face {face_ms:.4f} update {upd_ms:.4f}")

    out = {
        "metric": "FP64 cell-stage updates/s",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "n_cells": n,
            "order": wl.meta.get("order"),
            "mesh": wl.meta,
            "rk": "SSP3",
            "dt": dt,
            "parallelism": "single-gpu",
        },
        "kernels_ms": {"reconstruct": rec_ms, "face_flux": face_ms, "update": upd_ms},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_reconstruct",
            "bytes_per_cell": info.recon_bytes_per_cell,
        },
        "setup_s": t_setup,
        "measured_copy_gbs": z.bandwidth_probe(1 << 30, 10),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
