"""ctypes binding of the test-only CPU emulation (tests/emu/libzfv_emu.so):
the product's setup + kernel math run in host loops (see tests/emu/emu.cpp)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "emu", "build", "libzfv_emu.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    subprocess.run(["make", "-C", os.path.join(_HERE, "emu")], check=True, stdout=subprocess.DEVNULL)
    from zisafvm_amd import CMesh, Params

    lib = ctypes.CDLL(LIB)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.zfv_emu_create.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(Params), ctypes.POINTER(vp)]
    lib.zfv_emu_destroy.argtypes = [vp]
    lib.zfv_emu_rate.argtypes = [vp, dp, dp]
    lib.zfv_emu_stencils.argtypes = [vp, ctypes.c_int, ip, ctypes.c_int, ip, ip, ctypes.c_int]
    lib.zfv_emu_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), dp]
    lib.zfv_emu_validate.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    lib.zfv_emu_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


class Emu:
    def __init__(self, mesh, params):
        self.lib = load()
        self.mesh = mesh
        self._cm = mesh.c_struct()
        h = ctypes.c_void_p()
        if self.lib.zfv_emu_create(ctypes.byref(self._cm), ctypes.byref(params), ctypes.byref(h)) != 0:
            raise RuntimeError(self.lib.zfv_emu_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.zfv_emu_destroy(self.h)

    def rate(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        dp = ctypes.POINTER(ctypes.c_double)
        if self.lib.zfv_emu_rate(self.h, u.ctypes.data_as(dp), out.ctypes.data_as(dp)) != 0:
            raise RuntimeError(self.lib.zfv_emu_last_error().decode())
        return out

    def stencils(self, i):
        l2g = (ctypes.c_int32 * 512)()
        size = ctypes.c_int32()
        mem = (ctypes.c_int32 * 1024)()
        w = self.lib.zfv_emu_stencils(self.h, i, l2g, 512, ctypes.byref(size), mem, 1024)
        assert w >= 0
        return list(l2g[: size.value]), list(mem[:w])

    def validate(self):
        """None when every index the kernels derive from the setup data is in
        range, else the first violation."""
        buf = ctypes.create_string_buffer(256)
        return None if self.lib.zfv_emu_validate(self.h, buf, 256) == 0 else buf.value.decode()

    def info(self):
        n, ne, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        self.lib.zfv_emu_info(self.h, ctypes.byref(n), ctypes.byref(ne), ctypes.byref(c))
        return n.value, ne.value, c.value


class EmuBatch:
    """The batch kernels' per-block functions over the library's block tables,
    in host loops (zfv_emu_batch_* in tests/emu/emu.cpp)."""

    def __init__(self, members, workloads, cfls, t_ends, max_steps):
        self.lib = load()
        lib = self.lib
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        lib.zfv_emu_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(dp), ctypes.POINTER(dp),
                                             dp, dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(vp)]
        lib.zfv_emu_batch_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.zfv_emu_batch_member.argtypes = [vp, ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]
        lib.zfv_emu_batch_counts.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int64)] * 3
        lib.zfv_emu_batch_destroy.argtypes = [vp]
        n = len(members)
        self.members, self.n_cells = members, [w.mesh.n_cells for w in workloads]
        self._u0 = [np.ascontiguousarray(w.u0, dtype=np.float64) for w in workloads]
        self._st = [None if w.steady is None else np.ascontiguousarray(w.steady, dtype=np.float64) for w in workloads]
        hs = (vp * n)(*[m.h for m in members])
        u0p = (dp * n)(*[u.ctypes.data_as(dp) for u in self._u0])
        stp = (dp * n)(*[None if s is None else s.ctypes.data_as(dp) for s in self._st])
        c = np.ascontiguousarray(cfls, dtype=np.float64)
        te = np.ascontiguousarray(t_ends, dtype=np.float64)
        ms = (ctypes.c_int * n)(*max_steps)
        h = vp()
        if lib.zfv_emu_batch_create(hs, n, u0p, stp, c.ctypes.data_as(dp), te.ctypes.data_as(dp), ms,
                                    ctypes.byref(h)) != 0:
            raise RuntimeError(lib.zfv_emu_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.zfv_emu_batch_destroy(self.h)

    def run(self, n_steps, waves=1, xcd_chunks=True, rec_sample=1):
        if self.lib.zfv_emu_batch_run(self.h, n_steps, waves, int(xcd_chunks), rec_sample) != 0:
            raise RuntimeError(self.lib.zfv_emu_last_error().decode())

    def member(self, k):
        u = np.zeros((self.n_cells[k], 5))
        t = ctypes.c_double()
        steps, active = ctypes.c_int(), ctypes.c_int()
        dp = ctypes.POINTER(ctypes.c_double)
        self.lib.zfv_emu_batch_member(self.h, k, u.ctypes.data_as(dp), ctypes.byref(t), ctypes.byref(steps),
                                      ctypes.byref(active))
        return u, t.value, steps.value, bool(active.value)

    def counts(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.lib.zfv_emu_batch_counts(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value
