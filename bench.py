"""Benchmark: FP64 cell-stage updates/s of the SSP3 rate-of-change hot path on
one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload: the 2D well-balanced configuration — isothermal atmosphere
rho = p = exp(-y) with a 1e-6 Gaussian pressure bump on a seeded jittered
360 x 360 Delaunay lattice (259,200 triangles inside the unit square plus
frozen ghost layers), CWENO-AO order 4, HLLC, SSP3, CFL 0.45.

One step = one SSP3 step (3 stages) of the whole instance, replayed from a
captured HIP graph with the state resident in HBM.  value = n_interior * 3 * K
/ T over exactly K timed steps, counting the interior cells (the
configuration's triangles or tetrahedra; the frozen ghost layers are not
counted).  The JSON line also carries
  - roofline: the dominant kernel (k_reconstruct), algorithmic bytes per launch
    / its average HIP-event duration on the handle's stream;
  - cpu_baseline: the oracle (the reference's OpenMP CPU path restated) timed
    on all of this host's cores over a bounded sample, and on one thread.
Small legs, run by default, each in a child process (its own HIP context;
a failure there is reported in its entry and leaves this run's line intact),
at the member sizes the batch and serving tests run:
  batched_small  batched_vs_plain on 8 members of the batched configuration's
                 mix at 1,000-2,000 triangles (--no-small skips both);
  serving_small  Poisson streams of 16 small requests (2D 1,000-2,000
                 triangles, 3D 3,200-5,000 tetrahedra) through the serving
                 runtime: p50/p99 latency.
Opt-in legs (the library refuses batches above 65,536 cells or 8 members,
so at the configurations' sizes both legs report that refusal):
  --batch  batched_vs_plain: the batched configuration (64 members, 5k-200k
           cells, order 3, per-member CFL every step, one graph per step for
           all members) against the same members stepped one after another on
           the plain path;
  --serve  serving: Poisson request streams through the serving runtime,
           p50/p99 latency and aggregate cell-stage updates/s.
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
METRIC = "FP64 cell-stage updates/s, batched vs plain, 1 GPU; p50/p99 solve latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_cores():
    """Host threads this process may run on (its affinity), all of them."""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def build_workload(name, n):
    from zisafvm_amd import workloads as W

    if name == "vortex2d":
        return W.vortex_2d(n=n or 128)
    if name == "isothermal2d":
        return W.isothermal_2d(n=n or 360)
    if name == "polytrope3d":
        return W.polytrope_3d(n=n or 40)
    raise SystemExit(f"unknown workload {name}")


def measured_traffic(workload, n_cells):
    """HBM bytes per k_reconstruct launch from the committed PMC summary of
    this workload (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter
    reads for a copy of known size), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)[f"k_reconstruct|{workload}|{n_cells}"]
        return t["fetch_bytes"] + t["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(wl, budget_s, n_counted):
    """Oracle rate-of-change calls (one RK stage each) on all of this host's
    cores for about budget_s, then on one thread for a shorter sample;
    cell-stage updates/s counted over the same cells as the GPU value."""
    from oracle.zfv_ref import Oracle

    cores = cpu_cores()
    o = Oracle(wl.mesh, wl.params, threads=cores)

    def timed(threads, budget, max_calls):
        o.set_threads(threads)
        t0 = time.perf_counter()
        o.rate(wl.u0)
        first = time.perf_counter() - t0
        calls = int(min(max_calls, max(1, budget / max(first, 1e-3))))
        t0 = time.perf_counter()
        for _ in range(calls):
            o.rate(wl.u0)
        return calls, time.perf_counter() - t0

    calls, el = timed(cores, budget_s, 20)
    calls1, el1 = timed(1, budget_s / 3.0, 5)
    n = wl.mesh.n_cells
    return {
        "value": n_counted * calls / el,
        "unit": "cell-stage updates/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": cpu_model(),
        "single_thread_value": n_counted * calls1 / el1,
        "sample": f"{calls} rate-of-change evaluations (one RK stage each) of the {wl.name} instance "
        f"({n} cells, {n_counted} counted) by the oracle, OpenMP on all {cores} host threads, {el:.1f} s; "
        f"single thread: {calls1} evaluations, {el1:.1f} s",
    }


def run_single(z, wl, steps, warmup):
    t0 = time.perf_counter()
    s = z.Solver(wl.mesh, wl.params)
    t_setup = time.perf_counter() - t0
    s.set_state(wl.u0)
    if wl.steady is not None:
        s.set_steady_state(wl.steady)
    info = s.info()
    dt = s.cfl_dt(wl.cfl)
    s.steps(warmup, 0.0, dt)
    s.synchronize()
    t0 = time.perf_counter()
    s.steps(steps, 0.0, dt)
    s.synchronize()
    el = time.perf_counter() - t0
    return s, info, dt, el, t_setup


def batched_vs_plain(z, steps, wls):
    """Members of a batched workload: one graph per step for all of them,
    versus the same members one after another on the plain path."""
    cells = sum(w.mesh.n_cells for w in wls)
    try:
        b = z.Batch(wls[0].params, cell_capacity=cells, face_capacity=2 * cells, max_instances=len(wls))
        b.close()
    except z.ZfvError as e:
        return {"members": len(wls), "cells": cells, "refused": str(e)}
    t0 = time.perf_counter()
    solvers = []
    for w in wls:
        s = z.Solver(w.mesh, w.params)
        s.set_state(w.u0)
        if w.steady is not None:
            s.set_steady_state(w.steady)
        solvers.append(s)
    t_setup = time.perf_counter() - t0
    faces = sum(s.info().n_faces for s in solvers)
    # plain: each member alone, graph-replayed steps with its own CFL dt
    dts = [s.cfl_dt(w.cfl) for s, w in zip(solvers, wls)]
    for s, dt in zip(solvers, dts):
        s.steps(1, 0.0, dt)
    t0 = time.perf_counter()
    for s, dt in zip(solvers, dts):
        s.steps(steps, 0.0, dt)
    el_plain = time.perf_counter() - t0
    # batched: all members in one batch, per-member CFL every step
    for s, w in zip(solvers, wls):
        s.set_state(w.u0)
    b = z.Batch(wls[0].params, cell_capacity=cells, face_capacity=faces, max_instances=len(wls))
    for s, w in zip(solvers, wls):
        b.add(s, w.cfl, 0.0, 1e9, 0)
    b.run(1)
    tuning = b.tune(-1).as_dict()
    t0 = time.perf_counter()
    b.run(steps)
    el_batch = time.perf_counter() - t0
    plain = cells * 3 * steps / el_plain
    batched = cells * 3 * steps / el_batch
    b.close()
    return {
        "members": len(wls),
        "cells": cells,
        "steps": steps,
        "plain_cell_stage_per_s": plain,
        "batched_cell_stage_per_s": batched,
        "batched_over_plain": batched / plain,
        "plain_ms_per_step_all_members": 1e3 * el_plain / steps,
        "batched_ms_per_step": 1e3 * el_batch / steps,
        "setup_s": t_setup,
        "tuning": tuning,
        "note": "batched includes the per-member CFL reduction every step; plain replays fixed-dt graphs",
    }


def serving(z, rates, n_requests, budget_gib):
    """Poisson streams of mixed 2D/3D solve requests through the serving
    runtime (continuous batching, HBM-budget admission); one pass per rate."""
    from zisafvm_amd import serving as S
    from zisafvm_amd import workloads as W

    try:
        b = z.Batch(z.make_params(), cell_capacity=1 << 23, face_capacity=1 << 24, max_instances=64)
        b.close()
    except z.ZfvError as e:
        return {"requests": n_requests, "refused": str(e)}
    t0 = time.perf_counter()
    reqs = W.serving_requests(n_requests)
    t_gen = time.perf_counter() - t0
    budget = int(budget_gib * (1 << 30))
    est = [S.estimate_device_bytes(w.mesh, w.params) for w in reqs]
    runs = []
    for rate in rates:
        srv = S.Server(budget, batch_cell_capacity=1 << 23, max_members=64, setup_workers=2)
        res = S.poisson_run(srv, reqs, rate, seed=int(rate * 7))
        srv.close()
        runs.append(res.summary())
        log(f"serving rate {rate}/s: {json.dumps(runs[-1])}")
    return {
        "requests": n_requests,
        "mix": "every 4th a 3D polytrope (10k-60k tets), others 2D vortex / isothermal (10k-200k triangles), "
               "t_end U[0.05, 0.5], CWENO-AO 3",
        "cells": [int(w.mesh.n_cells) for w in reqs],
        "hbm_budget_bytes": budget,
        "max_request_estimate_bytes": int(max(est)),
        "request_generation_s": t_gen,
        "latency_definition": "submit to result, including host setup of the instance",
        "runs": runs,
    }


def serving_small(z, rates):
    """Poisson streams of small requests (the serving tests' sizes) through
    the serving runtime: p50/p99 latency per arrival rate."""
    from zisafvm_amd import serving as S
    from zisafvm_amd import workloads as W

    reqs = W.serving_requests(16, min_tris=1000, max_tris=2000, min_tets=3200, max_tets=5000, t_min=0.005,
                              t_max=0.05)
    est = [S.estimate_device_bytes(w.mesh, w.params) for w in reqs]
    runs = []
    for rate in rates:
        srv = S.Server(4 * max(est), batch_cell_capacity=1 << 16, max_members=8, setup_workers=2)
        res = S.poisson_run(srv, reqs, rate, seed=int(rate * 7))
        srv.close()
        runs.append(res.summary())
    return {
        "requests": len(reqs),
        "mix": "every 4th a 3D polytrope (3,200-5,000 tets), others 2D vortex / isothermal (1,000-2,000 "
               "triangles), t_end U[0.005, 0.05], CWENO-AO 3 (not the serving configuration's sizes)",
        "cells": [int(w.mesh.n_cells) for w in reqs],
        "hbm_budget_bytes": int(4 * max(est)),
        "latency_definition": "submit to result, including host setup of the instance",
        "runs": runs,
    }


def child_leg(name, timeout_s=240):
    """Runs one small leg in a child process and returns its JSON (or what
    went wrong); this process's own results are unaffected either way."""
    import subprocess

    try:
        r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--leg", name], capture_output=True,
                           text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s} s"}
    for line in reversed(r.stdout.splitlines()):
        if line.startswith("{"):
            return json.loads(line)
    return {"error": f"exit status {r.returncode}", "stderr_tail": r.stderr[-800:]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="isothermal2d")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", action="store_true",
                    help="add the batched-vs-plain leg (64 members)")
    ap.add_argument("--batch-small", action="store_true",
                    help="add a batched-vs-plain leg of 8 small members (1,000-2,000 triangles)")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--serve", action="store_true", help="add the serving leg (Poisson request streams)")
    ap.add_argument("--serve-rates", default="2,10,50", help="Poisson arrival rates (requests/s)")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-budget-gib", type=float, default=1.0)
    ap.add_argument("--no-small", action="store_true", help="skip the small batched and serving legs")
    ap.add_argument("--leg", default="", help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run only the streaming-copy calibration kernel (1 GiB x 5) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): it measures one GPU's share of the "
              f"work, and no multi-GPU figure is produced.", file=sys.stderr)
        return 2

    import zisafvm_amd as z

    if args.leg == "batch-small":
        from zisafvm_amd import workloads as W

        res = batched_vs_plain(z, 200, W.batched_2d(n_instances=8, min_cells=1000, max_cells=2000))
        res["workload"] = "8 members of 1,000-2,000 triangles (not the 64-member configuration)"
        print(json.dumps(res), flush=True)
        return 0
    if args.leg == "serve-small":
        print(json.dumps(serving_small(z, [50.0, 200.0, 1000.0])), flush=True)
        return 0

    if args.copy_probe_only:
        bw = z.bandwidth_probe(1 << 30, 5)
        print(json.dumps({"copy_probe_bytes": 1 << 30, "iters": 5, "gbytes_per_s": bw}), flush=True)
        return 0

    t0 = time.perf_counter()
    wl = build_workload(args.workload, args.n)
    log(f"workload {wl.name}: {wl.mesh.n_cells} cells, mesh+IC {time.perf_counter() - t0:.1f}s")
    s, info, dt, el, t_setup = run_single(z, wl, args.steps, args.warmup)
    n = wl.mesh.n_cells
    # cell-stage updates are counted over the interior cells, whose states
    # the steps advance; ghost cells are held at the steady state by the
    # frozen BC (those away from the interior are not even reconstructed)
    n_interior = int((wl.mesh.ghost == 0).sum())
    value = n_interior * 3 * args.steps / el
    log(f"setup {t_setup:.1f}s; {args.steps} steps in {el:.3f}s: {value:.3e} cell-stage updates/s "
        f"({n_interior} interior cells of {n})")
    # per-kernel HIP-event times right after the timed steps, with the GPU
    # still at the clocks the steps ran at
    kt = s.time_kernels(max(2, min(10, args.steps)), dt)

    # the device time loop (zfv_advance): LocalCFL chosen on the device every
    # step, dt clamped to the end time, no host round trip per step
    # untimed first call: captures and instantiates the step graph
    s.advance(wl.cfl, 0.0, 1.0e9, max_steps=max(1, args.warmup), check_every=args.steps)
    s.synchronize()
    t0 = time.perf_counter()
    _, adv_steps = s.advance(wl.cfl, 0.0, 1.0e9, max_steps=args.steps, check_every=args.steps)
    el_adv = time.perf_counter() - t0
    # the same fixed-dt steps with a snapshot of the state draining to the
    # host on a second stream behind them (zfv_snapshot_begin/_end)
    s.snapshot_begin()
    s.snapshot_end()
    s.synchronize()
    t0 = time.perf_counter()
    s.snapshot_begin()
    s.steps(args.steps, 0.0, dt)
    t1 = time.perf_counter()
    s.snapshot_end()
    el_snap, snap_end_ms = t1 - t0, 1e3 * (time.perf_counter() - t1)
    # the drop-in RateOfChange::compute path: host state in, host tendency out
    # (AoS <-> SoA on the host and PCIe both ways); reported beside, never value
    dudt = wl.u0 * 0.0
    s.rate(wl.u0, dudt=dudt)
    t0 = time.perf_counter()
    n_rate = 10
    for _ in range(n_rate):
        s.rate(wl.u0, dudt=dudt)
    rate_ms = 1e3 * (time.perf_counter() - t0) / n_rate
    launches = kt.n_launches
    rec_ms = kt.reconstruct_ms / launches
    face_ms = kt.face_flux_ms / launches
    upd_ms = kt.update_ms / launches
    # cells the reconstruction kernel processes: under the frozen BC (a steady
    # state is set) ghost cells with no interior neighbour are skipped
    n_rec = n - info.n_far_ghost_cells if wl.steady is not None else n
    achieved = info.recon_bytes_per_cell * n_rec / (rec_ms * 1e-3) / 1e9
    log(f"kernel avg ms: reconstruct {rec_ms:.4f} face {face_ms:.4f} update {upd_ms:.4f}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "n_cells": n,
            "n_interior_cells": n_interior,
            "value_counts": "interior cells (the configuration's triangles / tetrahedra); ghost layers excluded",
            "n_reconstructed_cells": n_rec,
            "order": wl.meta.get("order"),
            "mesh": wl.meta,
            "reconstruction": "CWENO-AO",
            "flux": "HLLC" if wl.params.flux == 0 else "Rusanov",
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
            "traffic": measured_traffic(wl.name, n),
            "cells_per_launch": n_rec,
            "kernel": "k_reconstruct",
            "bytes_per_cell": info.recon_bytes_per_cell,
            "traffic_source": "copied from profiles/traffic.json, not measured in this run: rocprofv3 "
                              "FETCH_SIZE/WRITE_SIZE of this workload's k_reconstruct, each scaled by what the "
                              "counter reads for a copy kernel of known size",
        },
        "setup_s": t_setup,
        "device_time_loop": {"steps": adv_steps, "ms_per_step": 1e3 * el_adv / max(adv_steps, 1),
                             "cell_stage_per_s": n_interior * 3 * adv_steps / el_adv,
                             "note": "zfv_advance: CFL dt chosen on the device every step"},
        "snapshot_overlap": {"ms_per_step": 1e3 * el_snap / args.steps, "state_mb": n * 5 * 8 / 1e6,
                             "end_ms": snap_end_ms,
                             "note": "the fixed-dt steps with a snapshot copied to the host behind them; "
                                     "end_ms: zfv_snapshot_end after them (wait + reorder on the host)"},
        "host_buffer_rate": {"ms_per_call": rate_ms, "cell_stage_per_s": n_interior / (rate_ms * 1e-3),
                             "note": "zfv_rate on host buffers (one RoC evaluation incl. PCIe and layout "
                                     "conversion), the drop-in RateOfChange::compute path"},
        "tuning": s.tune(-1).as_dict(),
        "measured_copy_gbs": z.bandwidth_probe(1 << 30, 10),
    }
    s.close()
    if not args.no_small:
        log("small batched and serving legs (child processes) ...")
        out["batched_small"] = child_leg("batch-small")
        log(json.dumps(out["batched_small"]))
        out["serving_small"] = child_leg("serve-small")
        log(json.dumps(out["serving_small"]))
    if args.batch or args.batch_small:
        from zisafvm_amd import workloads as W
    if args.batch:
        log("batched vs plain ...")
        out["batched_vs_plain"] = batched_vs_plain(z, args.batch_steps, W.batched_2d())
        log(json.dumps(out["batched_vs_plain"]))
    if args.batch_small:
        # the batched configuration's mix (vortices and atmospheres, order 3)
        # at the member sizes the batch tests run: 8 members of 1,000-2,000
        # triangles, where per-member launches dominate the plain path
        log("batched vs plain, small members ...")
        wls = W.batched_2d(n_instances=8, min_cells=1000, max_cells=2000)
        out["batched_small"] = batched_vs_plain(z, args.batch_steps, wls)
        out["batched_small"]["workload"] = "8 members of 1,000-2,000 triangles (not the 64-member configuration)"
        log(json.dumps(out["batched_small"]))
    if args.serve:
        log("serving ...")
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        out["serving"] = serving(z, rates, args.serve_requests, args.serve_budget_gib)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_budget, n_interior)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
