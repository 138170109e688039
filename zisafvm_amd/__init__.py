"""zisafvm_amd — host-side (Python) access to libzfv_hip.so, the gfx950 FP64
rate-of-change engine for unstructured CWENO-AO/WENO-AO finite volumes.

The C-ABI is declared in ``include/zfv.h``; this module binds it with ctypes
for tests and benchmarks.  The hot path is the HIP library: nothing here
computes a rate of change, and a missing library raises instead of falling
back to anything else.

Reference interface mirrored (ZisaFVM): ``RateOfChange::compute``
(include/zisa/ode/rate_of_change.hpp:37-39) -> :meth:`Solver.rate`;
``TimeIntegration::compute_step`` (include/zisa/ode/time_integration.hpp:42-44)
-> :meth:`Solver.step`; ``LocalCFL::operator()``
(include/zisa/model/local_cfl_condition_impl.hpp:25-40) -> :meth:`Solver.cfl_dt`.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libzfv_hip.so")

ZFV_CWENO_AO, ZFV_WENO_AO = 0, 1
ZFV_FLUX_HLLC, ZFV_FLUX_RUSANOV = 0, 1
ZFV_EQ_NONE, ZFV_EQ_ISENTROPIC, ZFV_EQ_ISOTHERMAL = 0, 1, 2
ZFV_GRAVITY_NONE, ZFV_GRAVITY_AXIAL, ZFV_GRAVITY_RADIAL, ZFV_GRAVITY_POLYTROPE = 0, 1, 2, 3
ZFV_FLUX_BC_NONE, ZFV_FLUX_BC_CONSTANT, ZFV_FLUX_BC_ISENTROPIC = 0, 1, 2


class Params(ctypes.Structure):
    """``zfv_params`` (include/zfv.h)."""

    _fields_ = [
        ("n_dims", ctypes.c_int32),
        ("n_stencils", ctypes.c_int32),
        ("orders", ctypes.c_int32 * 5),
        ("biases", ctypes.c_int32 * 5),
        ("overfit", ctypes.c_double * 5),
        ("linear_weights", ctypes.c_double * 5),
        ("epsilon", ctypes.c_double),
        ("exponent", ctypes.c_double),
        ("reconstruction", ctypes.c_int32),
        ("flux", ctypes.c_int32),
        ("equilibrium", ctypes.c_int32),
        ("gravity", ctypes.c_int32),
        ("gravity_g", ctypes.c_double),
        ("gravity_axis", ctypes.c_double * 3),
        ("polytrope_rhoC", ctypes.c_double),
        ("polytrope_K", ctypes.c_double),
        ("polytrope_G", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("gas_constant", ctypes.c_double),
        ("edge_deg", ctypes.c_int32),
        ("volume_deg", ctypes.c_int32),
        ("moments_deg", ctypes.c_int32),
        ("flux_bc", ctypes.c_int32),
        ("steps_per_recompute", ctypes.c_int32),
        ("scaling", ctypes.c_int32),
        ("recompute_threshold", ctypes.c_double),
    ]


class CMesh(ctypes.Structure):
    """``zfv_mesh`` (include/zfv.h)."""

    _fields_ = [
        ("n_dims", ctypes.c_int32),
        ("n_cells", ctypes.c_int32),
        ("vertex_ids", ctypes.POINTER(ctypes.c_int32)),
        ("corners", ctypes.POINTER(ctypes.c_double)),
        ("ghost_mask", ctypes.POINTER(ctypes.c_uint8)),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int64),
        ("n_faces", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("bytes_per_cell_stage", ctypes.c_double),
        ("recon_bytes_per_cell", ctypes.c_double),
        ("max_stencil_size", ctypes.c_int32),
        ("n_points_face", ctypes.c_int32),
        ("n_far_ghost_cells", ctypes.c_int64),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("rec_threads", ctypes.c_int32),
        ("cells_per_block", ctypes.c_int32),
        ("tuned", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("rec_ms", ctypes.c_double * 3),
    ]

    def as_dict(self):
        return {"rec_threads": self.rec_threads, "cells_per_block": self.cells_per_block,
                "rec_ms_at_64_128_256_threads": list(self.rec_ms)}


class KernelTimes(ctypes.Structure):
    _fields_ = [
        ("reconstruct_ms", ctypes.c_double),
        ("face_flux_ms", ctypes.c_double),
        ("update_ms", ctypes.c_double),
        ("n_launches", ctypes.c_int32),
    ]


def make_params(
    n_dims=2,
    order=3,
    n_biased=None,
    overfit_central=2.0,
    overfit_biased=1.5,
    linear_weights=(100.0, 1.0, 1.0, 1.0, 1.0),
    epsilon=1e-6,
    exponent=4.0,
    reconstruction=ZFV_CWENO_AO,
    flux=ZFV_FLUX_HLLC,
    equilibrium=ZFV_EQ_NONE,
    gravity=ZFV_GRAVITY_NONE,
    gravity_g=0.0,
    gravity_axis=(0.0, 1.0, 0.0),
    polytrope=(1.0, 1.0, 1.0),
    gamma=1.4,
    gas_constant=1.0,
    edge_deg=None,
    volume_deg=None,
    moments_deg=-1,
    flux_bc=0,
    steps_per_recompute=1,
    recompute_threshold=1e300,
) -> Params:
    """Reconstruction parameters in the reference's shape
    (src/zisa/reconstruction/hybrid_weno_params.cpp:26-44: orders {p,2,2,2},
    biases {c,b,b,b}, overfit {2,1.5,1.5,1.5}, linear weights {100,1,1,1},
    epsilon 1e-6, exponent 4); one one-sided stencil per face by default
    (three on triangles, four on tetrahedra)."""
    p = Params()
    nb = (n_dims + 1) if n_biased is None else n_biased
    p.n_dims = n_dims
    p.n_stencils = 1 + nb
    for k in range(5):
        p.orders[k] = order if k == 0 else 2
        p.biases[k] = 0 if k == 0 else 1
        p.overfit[k] = overfit_central if k == 0 else overfit_biased
        p.linear_weights[k] = linear_weights[k] if k < len(linear_weights) else 1.0
    p.epsilon = epsilon
    p.exponent = exponent
    p.reconstruction = reconstruction
    p.flux = flux
    p.equilibrium = equilibrium
    p.gravity = gravity
    p.gravity_g = gravity_g
    for d in range(3):
        p.gravity_axis[d] = gravity_axis[d]
    p.polytrope_rhoC, p.polytrope_K, p.polytrope_G = polytrope
    p.gamma = gamma
    p.gas_constant = gas_constant
    if edge_deg is None:
        edge_deg = order if n_dims == 2 else 2
    if volume_deg is None:
        volume_deg = 4 if n_dims == 2 else 3
    p.edge_deg = edge_deg
    p.volume_deg = volume_deg
    p.moments_deg = moments_deg
    p.flux_bc = flux_bc
    p.steps_per_recompute = steps_per_recompute
    p.recompute_threshold = recompute_threshold
    return p


class ZfvError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libzfv_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ZfvError(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    c_int, c_double, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    sigs = {
        "zfv_create": (c_int, [ctypes.POINTER(CMesh), ctypes.POINTER(Params), ctypes.POINTER(vp)]),
        "zfv_destroy": (c_int, [vp]),
        "zfv_set_state": (c_int, [vp, dp]),
        "zfv_get_state": (c_int, [vp, dp]),
        "zfv_set_steady_state": (c_int, [vp, dp]),
        "zfv_rate": (c_int, [vp, dp, dp, c_double]),
        "zfv_step": (c_int, [vp, c_double, c_double]),
        "zfv_steps": (c_int, [vp, ctypes.c_int32, c_double, c_double]),
        "zfv_cfl_dt": (c_int, [vp, c_double, dp]),
        "zfv_advance": (c_int, [vp, c_double, c_double, c_double, ctypes.c_int32, ctypes.c_int32, c_double,
                                ctypes.c_int32, dp, ip, ip]),
        "zfv_advance_noalloc": (c_int, [vp, c_double, c_double, c_double, ctypes.c_int32, ctypes.c_int32, c_double,
                                        ctypes.c_int32, dp, ip, ip]),
        "zfv_synchronize": (c_int, [vp]),
        "zfv_snapshot_begin": (c_int, [vp]),
        "zfv_snapshot_end": (c_int, [vp, dp]),
        "zfv_get_info": (c_int, [vp, ctypes.POINTER(Info)]),
        "zfv_get_stencils": (c_int, [vp, ctypes.c_int32, ip, ctypes.c_int32, ip, ip, ctypes.c_int32]),
        "zfv_time_kernels": (c_int, [vp, ctypes.c_int32, c_double, ctypes.POINTER(KernelTimes)]),
        "zfv_last_error": (ctypes.c_char_p, []),
        "zfv_bandwidth_probe": (c_int, [ctypes.c_int64, ctypes.c_int32, dp]),
        "zfv_batch_create": (c_int, [ctypes.POINTER(Params), ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                     ctypes.POINTER(vp)]),
        "zfv_batch_destroy": (c_int, [vp]),
        "zfv_batch_add": (c_int, [vp, vp, c_double, c_double, c_double, ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_int32)]),
        "zfv_batch_remove": (c_int, [vp, ctypes.c_int32]),
        "zfv_batch_run": (c_int, [vp, ctypes.c_int32]),
        "zfv_batch_launch": (c_int, [vp, ctypes.c_int32]),
        "zfv_batch_wait": (c_int, [vp]),
        "zfv_batch_query": (c_int, [vp, ctypes.c_int32, dp, dp, ip, ip]),
        "zfv_batch_set_xcd_chunks": (c_int, [vp, ctypes.c_int32]),
        "zfv_tune": (c_int, [vp, ctypes.c_int32, ctypes.POINTER(Tuning)]),
        "zfv_batch_tune": (c_int, [vp, ctypes.c_int32, ctypes.POINTER(Tuning)]),
        "zfv_mesh_lattice2d": (
            c_int,
            [ctypes.c_int32, c_double, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
             ctypes.POINTER(vp)],
        ),
        "zfv_mesh_ball3d": (
            c_int, [ctypes.c_int32, c_double, c_double, c_double, ctypes.c_uint64, ctypes.POINTER(vp)]
        ),
        "zfv_mesh_view": (c_int, [vp, ctypes.POINTER(CMesh)]),
        "zfv_mesh_free": (c_int, [vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = load_library().zfv_last_error()
        raise ZfvError(f"zfv error {rc}: {msg.decode() if msg else ''}")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Mesh:
    """A simplex mesh in the reference's layout plus per-cell corners."""

    n_dims: int
    vertex_ids: np.ndarray  # (n, d+1) int32
    corners: np.ndarray  # (n, d+1, 3) float64
    ghost: np.ndarray  # (n,) uint8

    @property
    def n_cells(self) -> int:
        return int(self.vertex_ids.shape[0])

    def c_struct(self) -> CMesh:
        self.vertex_ids = np.ascontiguousarray(self.vertex_ids, dtype=np.int32)
        self.corners = np.ascontiguousarray(self.corners, dtype=np.float64)
        self.ghost = np.ascontiguousarray(self.ghost, dtype=np.uint8)
        m = CMesh()
        m.n_dims = self.n_dims
        m.n_cells = self.n_cells
        m.vertex_ids = self.vertex_ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        m.corners = self.corners.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        m.ghost_mask = self.ghost.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        return m

    def centroids(self) -> np.ndarray:
        return self.corners.mean(axis=1)

    def volumes(self) -> np.ndarray:
        c = self.corners
        if self.n_dims == 2:
            a, b = c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]
            return 0.5 * np.abs(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
        a, b, d = c[:, 1] - c[:, 0], c[:, 2] - c[:, 0], c[:, 3] - c[:, 0]
        return np.abs(np.einsum("ij,ij->i", a, np.cross(b, d))) / 6.0


def _mesh_from_buf(lib, buf) -> Mesh:
    view = CMesh()
    _check(lib.zfv_mesh_view(buf, ctypes.byref(view)))
    n, d = view.n_cells, view.n_dims
    vid = np.ctypeslib.as_array(view.vertex_ids, shape=(n * (d + 1),)).copy().reshape(n, d + 1)
    crn = np.ctypeslib.as_array(view.corners, shape=(n * (d + 1) * 3,)).copy().reshape(n, d + 1, 3)
    gh = np.ctypeslib.as_array(view.ghost_mask, shape=(n,)).copy()
    lib.zfv_mesh_free(buf)
    return Mesh(d, vid, crn, gh)


def bandwidth_probe(nbytes: int = 1 << 30, iters: int = 20) -> float:
    """Measured device copy rate (GB/s, read + write) of a plain streaming copy."""
    lib = load_library()
    r = ctypes.c_double()
    _check(lib.zfv_bandwidth_probe(nbytes, iters, ctypes.byref(r)))
    return r.value


def lattice_mesh_2d(n: int, jitter: float = 0.2, seed: int = 1, periodic=(True, True), ghost_layers: int = 3) -> Mesh:
    """Seeded jittered n x n lattice on [0,1]^2, Delaunay diagonals (2 n^2 cells
    per periodic domain; non-periodic directions get ghost layers)."""
    lib = load_library()
    buf = ctypes.c_void_p()
    _check(lib.zfv_mesh_lattice2d(n, jitter, seed, int(periodic[0]), int(periodic[1]), ghost_layers, ctypes.byref(buf)))
    return _mesh_from_buf(lib, buf)


def ball_mesh_3d(n: int, radius: float = 1.0, ghost_radius: float = 0.85, jitter: float = 0.1, seed: int = 1) -> Mesh:
    lib = load_library()
    buf = ctypes.c_void_p()
    _check(lib.zfv_mesh_ball3d(n, radius, ghost_radius, jitter, seed, ctypes.byref(buf)))
    return _mesh_from_buf(lib, buf)


class Solver:
    """One device-resident instance (a ``zfv_handle``)."""

    def __init__(self, mesh: Mesh, params: Params):
        self.lib = load_library()
        self.mesh = mesh
        self.params = params
        self._cm = mesh.c_struct()
        h = ctypes.c_void_p()
        _check(self.lib.zfv_create(ctypes.byref(self._cm), ctypes.byref(params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.zfv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_cells(self) -> int:
        return self.mesh.n_cells

    def set_state(self, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == (self.n_cells, 5)
        _check(self.lib.zfv_set_state(self.h, _dp(u)))

    def get_state(self) -> np.ndarray:
        u = np.zeros((self.n_cells, 5))
        _check(self.lib.zfv_get_state(self.h, _dp(u)))
        return u

    def snapshot_begin(self):
        """Starts copying the current device state off the device while later
        steps proceed (zfv_snapshot_begin)."""
        _check(self.lib.zfv_snapshot_begin(self.h))

    def snapshot_end(self) -> np.ndarray:
        """The state at the last snapshot_begin (zfv_snapshot_end)."""
        u = np.zeros((self.n_cells, 5))
        _check(self.lib.zfv_snapshot_end(self.h, _dp(u)))
        return u

    def set_steady_state(self, u: np.ndarray):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _check(self.lib.zfv_set_steady_state(self.h, _dp(u)))

    def rate(self, u: np.ndarray, t: float = 0.0, dudt: Optional[np.ndarray] = None) -> np.ndarray:
        """RateOfChange::compute: returns dudt + L(u) (dudt defaults to zeros)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u) if dudt is None else np.ascontiguousarray(dudt, dtype=np.float64)
        _check(self.lib.zfv_rate(self.h, _dp(u), _dp(out), t))
        return out

    def step(self, t: float, dt: float):
        _check(self.lib.zfv_step(self.h, t, dt))

    def steps(self, n: int, t: float, dt: float):
        _check(self.lib.zfv_steps(self.h, n, t, dt))

    def advance(self, cfl: float, t0: float, t_end: float, max_steps: int = 0, check_every: int = 16,
                drho_crit_rel: float = 0.0, max_exponent: int = 0):
        """TimeLoop on the device: LocalCFL steps from t0 to t_end (zfv_advance);
        returns (t_final, accepted steps).  drho_crit_rel > 0 rejects and
        retries steps as RejectLargeDensityChange(drho_crit_rel, max_exponent)
        does; the number of rejected steps is kept in self.rejected_steps."""
        t = ctypes.c_double(0.0)
        k = ctypes.c_int32(0)
        r = ctypes.c_int32(0)
        _check(self.lib.zfv_advance(self.h, cfl, t0, t_end, max_steps, check_every, drho_crit_rel, max_exponent,
                                    ctypes.byref(t), ctypes.byref(k), ctypes.byref(r)))
        self.rejected_steps = r.value
        return t.value, k.value

    def advance_noalloc(self, cfl: float, t0: float, t_end: float, max_steps: int = 0, check_every: int = 16,
                        drho_crit_rel: float = 0.0, max_exponent: int = 0):
        """The same time loop as advance() through zfv_advance_noalloc, which
        allocates nothing (its loop state lives in the library's device
        scratch array); returns (t_final, accepted steps)."""
        t = ctypes.c_double(0.0)
        k = ctypes.c_int32(0)
        r = ctypes.c_int32(0)
        _check(self.lib.zfv_advance_noalloc(self.h, cfl, t0, t_end, max_steps, check_every, drho_crit_rel,
                                            max_exponent, ctypes.byref(t), ctypes.byref(k), ctypes.byref(r)))
        self.rejected_steps = r.value
        return t.value, k.value

    def cfl_dt(self, cfl: float) -> float:
        dt = ctypes.c_double()
        _check(self.lib.zfv_cfl_dt(self.h, cfl, ctypes.byref(dt)))
        return dt.value

    def synchronize(self):
        _check(self.lib.zfv_synchronize(self.h))

    def info(self) -> Info:
        inf = Info()
        _check(self.lib.zfv_get_info(self.h, ctypes.byref(inf)))
        return inf

    def stencils(self, i: int):
        l2g = (ctypes.c_int32 * 512)()
        size = ctypes.c_int32()
        mem = (ctypes.c_int32 * 1024)()
        w = self.lib.zfv_get_stencils(self.h, i, l2g, 512, ctypes.byref(size), mem, 1024)
        if w < 0:
            _check(w)
        return list(l2g[: size.value]), list(mem[:w])

    def time_kernels(self, n_steps: int, dt: float) -> KernelTimes:
        kt = KernelTimes()
        _check(self.lib.zfv_time_kernels(self.h, n_steps, dt, ctypes.byref(kt)))
        return kt

    def tune(self, rec_threads: int = 0) -> Tuning:
        """Re-tune (0), report (-1) or fix (64/128/256 threads per block)
        the reconstruction launch configuration."""
        t = Tuning()
        _check(self.lib.zfv_tune(self.h, rec_threads, ctypes.byref(t)))
        return t


class Batch:
    """Many instances of one kernel shape advanced together (``zfv_batch``)."""

    def __init__(self, params: Params, cell_capacity: int, face_capacity: int, max_instances: int):
        self.lib = load_library()
        b = ctypes.c_void_p()
        _check(self.lib.zfv_batch_create(ctypes.byref(params), cell_capacity, face_capacity, max_instances,
                                         ctypes.byref(b)))
        self.b = b
        self.members = {}

    def close(self):
        if self.b:
            self.lib.zfv_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, solver: Solver, cfl: float, t0: float, t_end: float, max_steps: int = 0) -> int:
        slot = ctypes.c_int32()
        _check(self.lib.zfv_batch_add(self.b, solver.h, cfl, t0, t_end, max_steps, ctypes.byref(slot)))
        self.members[slot.value] = solver
        return slot.value

    def remove(self, slot: int):
        _check(self.lib.zfv_batch_remove(self.b, slot))
        self.members.pop(slot, None)

    def run(self, n_steps: int):
        _check(self.lib.zfv_batch_run(self.b, n_steps))

    def launch(self, n_steps: int):
        """Enqueue n_steps steps and return (the batch's own stream)."""
        _check(self.lib.zfv_batch_launch(self.b, n_steps))

    def wait(self):
        _check(self.lib.zfv_batch_wait(self.b))

    def query(self, slot: int):
        t, dt = ctypes.c_double(), ctypes.c_double()
        steps, active = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.zfv_batch_query(self.b, slot, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(steps),
                                        ctypes.byref(active)))
        return dict(t=t.value, dt=dt.value, steps=steps.value, active=bool(active.value))

    def set_xcd_chunks(self, on: bool):
        _check(self.lib.zfv_batch_set_xcd_chunks(self.b, int(on)))

    def tune(self, rec_threads: int = 0) -> Tuning:
        t = Tuning()
        _check(self.lib.zfv_batch_tune(self.b, rec_threads, ctypes.byref(t)))
        return t
