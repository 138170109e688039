"""The reference's mesh (.msh.h5) and snapshot / restart files, through
libzfv_io.so (include/zfv_io.h): HDF5 on the host, nothing on the device.

Mesh files carry n_dims, vertices and vertex_indices
(load_grid_gmsh_h5, src/zisa/grid/grid.cpp:889-901); snapshots carry time,
n_steps, one dataset per conserved variable (rho, mv1, mv2, mv3, E) and
n_avars (src/zisa/model/models.cpp:7-27, src/zisa/model/all_variables.cpp:74-113).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import CMesh, Mesh, ZfvError

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(_HERE, "_lib", "libzfv_io.so")
_io = None


def load_io_library(path: str = IO_LIB_PATH):
    """Load libzfv_io.so (raises if it is not built)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(path):
        raise ZfvError(f"{path} is not built (it needs the HDF5 headers); run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    c_int, vp = ctypes.c_int, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    lp = ctypes.POINTER(ctypes.c_int64)
    sigs = {
        "zfv_io_last_error": (ctypes.c_char_p, []),
        "zfv_io_read_msh_h5": (c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
        "zfv_io_mesh_view": (c_int, [vp, ctypes.POINTER(CMesh)]),
        "zfv_io_mesh_free": (c_int, [vp]),
        "zfv_io_write_msh_h5": (c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, dp, ctypes.c_int64, lp]),
        "zfv_io_renumber_msh_h5": (c_int, [ctypes.c_char_p, lp]),
        "zfv_io_write_snapshot": (c_int, [ctypes.c_char_p, ctypes.c_int64, dp, ctypes.c_double, ctypes.c_int64]),
        "zfv_io_write_steady_state": (c_int, [ctypes.c_char_p, ctypes.c_int64, dp]),
        "zfv_io_read_snapshot": (c_int, [ctypes.c_char_p, ctypes.c_int64, dp, dp, lp]),
        "zfv_io_snapshot_cells": (c_int, [ctypes.c_char_p, lp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _io = lib
    return lib


def _check(rc):
    if rc != 0:
        msg = load_io_library().zfv_io_last_error()
        raise ZfvError(f"zfv_io error {rc}: {msg.decode() if msg else ''}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def read_msh_h5(path: str) -> Mesh:
    """A .msh.h5 mesh as a Mesh (per-cell corners, no ghost cells)."""
    lib = load_io_library()
    buf = ctypes.c_void_p()
    _check(lib.zfv_io_read_msh_h5(os.fsencode(path), ctypes.byref(buf)))
    try:
        view = CMesh()
        lib.zfv_io_mesh_view(buf, ctypes.byref(view))
        n, d = view.n_cells, view.n_dims
        vid = np.ctypeslib.as_array(view.vertex_ids, shape=(n * (d + 1),)).copy().reshape(n, d + 1)
        crn = np.ctypeslib.as_array(view.corners, shape=(n * (d + 1) * 3,)).copy().reshape(n, d + 1, 3)
        gh = np.ctypeslib.as_array(view.ghost_mask, shape=(n,)).copy()
    finally:
        lib.zfv_io_mesh_free(buf)
    return Mesh(d, vid, crn, gh)


def write_msh_h5(path: str, n_dims: int, vertices: np.ndarray, vertex_indices: np.ndarray):
    vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    vertex_indices = np.ascontiguousarray(vertex_indices, dtype=np.int64).reshape(-1, n_dims + 1)
    _check(load_io_library().zfv_io_write_msh_h5(os.fsencode(path), n_dims, vertices.shape[0], _dp(vertices),
                                                 vertex_indices.shape[0], _lp(vertex_indices)))


def mesh_to_msh_h5(path: str, mesh: Mesh):
    """Writes a Mesh with vertices shared by id (the first corner seen for each
    vertex id gives its coordinates).  A .msh.h5 file has no periodic
    translations: meshes with unwrapped periodic corners do not round-trip."""
    vid = mesh.vertex_ids.astype(np.int64)
    nv = int(vid.max()) + 1
    verts = np.zeros((nv, 3))
    flat_ids = vid.ravel()
    flat_c = mesh.corners.reshape(-1, 3)
    first = np.unique(flat_ids, return_index=True)[1]
    verts[flat_ids[first]] = flat_c[first]
    write_msh_h5(path, mesh.n_dims, verts, vid)


def renumber_msh_h5(path: str) -> np.ndarray:
    """Hilbert renumbering of a .msh.h5 file in place (the reference's
    renumber_grid tool); returns the permutation (new cell k was old cell
    order[k])."""
    lib = load_io_library()
    n = ctypes.c_int64()
    # the cell count, to size the permutation
    buf = ctypes.c_void_p()
    _check(lib.zfv_io_read_msh_h5(os.fsencode(path), ctypes.byref(buf)))
    view = CMesh()
    lib.zfv_io_mesh_view(buf, ctypes.byref(view))
    n = view.n_cells
    lib.zfv_io_mesh_free(buf)
    order = np.zeros(n, dtype=np.int64)
    _check(lib.zfv_io_renumber_msh_h5(os.fsencode(path), _lp(order)))
    return order


def write_snapshot(path: str, cvars: np.ndarray, t: float, n_steps: int):
    """AllVariables snapshot (time, n_steps, rho, mv1, mv2, mv3, E, n_avars = 0)."""
    u = np.ascontiguousarray(cvars, dtype=np.float64).reshape(-1, 5)
    _check(load_io_library().zfv_io_write_snapshot(os.fsencode(path), u.shape[0], _dp(u), t, n_steps))


def write_steady_state(path: str, cvars: np.ndarray):
    u = np.ascontiguousarray(cvars, dtype=np.float64).reshape(-1, 5)
    _check(load_io_library().zfv_io_write_steady_state(os.fsencode(path), u.shape[0], _dp(u)))


def read_snapshot(path: str, with_time: bool = True):
    """Restart: (cvars n x 5, time, n_steps) of a snapshot; with_time False
    reads a steady state file (cvars only)."""
    lib = load_io_library()
    n = ctypes.c_int64()
    _check(lib.zfv_io_snapshot_cells(os.fsencode(path), ctypes.byref(n)))
    u = np.zeros((n.value, 5))
    t = ctypes.c_double()
    k = ctypes.c_int64()
    _check(lib.zfv_io_read_snapshot(os.fsencode(path), n.value, _dp(u), ctypes.byref(t) if with_time else None,
                                    ctypes.byref(k) if with_time else None))
    return (u, t.value, k.value) if with_time else u
