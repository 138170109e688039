"""Python side of the serving runtime (``zfv_server_*`` in include/zfv.h) and
the Poisson-arrival driver used by the serving benchmark.

The runtime itself is native (zisafvm_amd/csrc/server.cpp): setup workers,
HBM-budget admission and the continuous-batching GPU loop run in C++ threads;
this module only submits requests on a schedule and reads results back.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import CMesh, Params, _check, load_library

QUEUED, SETUP, RUNNING, DONE, FAILED = 0, 1, 2, 3, 4


class RequestStatus(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("n_cells", ctypes.c_int64),
        ("t_final", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("start_s", ctypes.c_double),
        ("latency_s", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("reserved_bytes", ctypes.c_int64),
        ("pending", ctypes.c_int32),
        ("in_setup", ctypes.c_int32),
        ("running", ctypes.c_int32),
        ("ready", ctypes.c_int32),
        ("cell_stages", ctypes.c_int64),
        ("steps_run", ctypes.c_int64),
    ]


def _bind(lib):
    if getattr(lib, "_zfv_server_bound", False):
        return
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    c_int = ctypes.c_int
    sig = {
        "zfv_server_create": (c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(vp)]),
        "zfv_server_submit": (c_int, [vp, ctypes.POINTER(CMesh), ctypes.POINTER(Params), dp, dp, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
        "zfv_server_poll": (c_int, [vp, ctypes.c_int64, ctypes.POINTER(RequestStatus)]),
        "zfv_server_result": (c_int, [vp, ctypes.c_int64, dp]),
        "zfv_server_release": (c_int, [vp, ctypes.c_int64]),
        "zfv_server_stats": (c_int, [vp, ctypes.POINTER(ServerStats)]),
        "zfv_server_destroy": (c_int, [vp]),
        "zfv_estimate_device_bytes": (c_int, [ctypes.POINTER(CMesh), ctypes.POINTER(Params),
                                              ctypes.POINTER(ctypes.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._zfv_server_bound = True


class Server:
    def __init__(self, hbm_budget_bytes: int, batch_cell_capacity: int = 1 << 24, max_members: int = 256,
                 setup_workers: int = 2):
        self.lib = load_library()
        _bind(self.lib)
        s = ctypes.c_void_p()
        _check(self.lib.zfv_server_create(hbm_budget_bytes, batch_cell_capacity, max_members, setup_workers,
                                          ctypes.byref(s)))
        self.s = s
        self._keep = {}

    def close(self):
        if self.s:
            self.lib.zfv_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, mesh, params, u0, steady=None, t_end=0.1, cfl=0.45, max_steps=0) -> int:
        cm = mesh.c_struct()
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        st = None if steady is None else np.ascontiguousarray(steady, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        rid = ctypes.c_int64()
        _check(self.lib.zfv_server_submit(self.s, ctypes.byref(cm), ctypes.byref(params), u0.ctypes.data_as(dp),
                                          None if st is None else st.ctypes.data_as(dp), t_end, cfl, max_steps,
                                          ctypes.byref(rid)))
        return rid.value

    def poll(self, rid: int) -> RequestStatus:
        st = RequestStatus()
        _check(self.lib.zfv_server_poll(self.s, rid, ctypes.byref(st)))
        return st

    def result(self, rid: int, n_cells: int) -> np.ndarray:
        u = np.zeros((n_cells, 5))
        _check(self.lib.zfv_server_result(self.s, rid, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return u

    def release(self, rid: int):
        self.lib.zfv_server_release(self.s, rid)

    def stats(self) -> ServerStats:
        st = ServerStats()
        self.lib.zfv_server_stats(self.s, ctypes.byref(st))
        return st

    def wait(self, rids, timeout_s: float = 600.0, poll_s: float = 0.002):
        t0 = time.perf_counter()
        pending = set(rids)
        while pending:
            for r in list(pending):
                if self.poll(r).state in (DONE, FAILED):
                    pending.discard(r)
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError(f"{len(pending)} requests unfinished after {timeout_s}s")
            if pending:
                time.sleep(poll_s)


def estimate_device_bytes(mesh, params) -> int:
    lib = load_library()
    _bind(lib)
    cm = mesh.c_struct()
    b = ctypes.c_int64()
    _check(lib.zfv_estimate_device_bytes(ctypes.byref(cm), ctypes.byref(params), ctypes.byref(b)))
    return b.value


@dataclass
class ServeResult:
    rate: float
    n_requests: int
    latencies: List[float]
    cell_stages: int
    wall_s: float

    def summary(self):
        lat = np.array(self.latencies)
        return {
            "arrival_rate_per_s": self.rate,
            "requests": self.n_requests,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "cell_stage_per_s": self.cell_stages / self.wall_s,
            "wall_s": self.wall_s,
        }


def poisson_run(server: Server, requests, rate: float, seed: int = 0, timeout_s: float = 900.0) -> ServeResult:
    """Submit `requests` (Workload objects with t_end set) with exponential
    inter-arrival times of mean 1/rate; wait for all; latency = result time
    - submit time as measured by the runtime."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(requests))
    t0 = time.perf_counter()
    next_t = 0.0
    rids = []
    for wl, gap in zip(requests, gaps):
        next_t += gap
        now = time.perf_counter() - t0
        if next_t > now:
            time.sleep(next_t - now)
        rids.append(server.submit(wl.mesh, wl.params, wl.u0, wl.steady, wl.t_end, wl.cfl,
                                  wl.n_steps or 0))
    server.wait(rids, timeout_s=timeout_s)
    wall = time.perf_counter() - t0
    lat, cs = [], 0
    for r in rids:
        st = server.poll(r)
        if st.state != DONE:
            try:
                server.result(r, st.n_cells)  # raises with the request's error
            except Exception as e:
                raise RuntimeError(f"request {r} failed: {e}") from None
            raise RuntimeError(f"request {r} failed")
        lat.append(st.latency_s)
        cs += st.n_cells * 3 * st.steps
        server.release(r)
    return ServeResult(rate, len(requests), lat, cs, wall)
