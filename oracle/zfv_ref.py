"""ctypes binding of the CPU oracle (oracle/build/libzfv_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product (zisafvm_amd)
never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libzfv_ref.so")


def build(force: bool = False) -> str:
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from zisafvm_amd import CMesh, Params  # struct layouts only (include/zfv.h)

    lib = ctypes.CDLL(LIB_PATH)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    c_int, c_double = ctypes.c_int, ctypes.c_double
    sig = {
        "zfv_ref_create": (c_int, [ctypes.POINTER(CMesh), ctypes.POINTER(Params), ctypes.POINTER(vp)]),
        "zfv_ref_destroy": (c_int, [vp]),
        "zfv_ref_rate": (c_int, [vp, dp, dp, c_double]),
        "zfv_ref_step": (c_int, [vp, dp, c_double, c_double]),
        "zfv_ref_cfl_dt": (c_int, [vp, dp, c_double, dp]),
        "zfv_ref_set_steady_state": (c_int, [vp, dp]),
        "zfv_ref_advance": (c_int, [vp, dp, c_double, c_double, c_double, c_int, c_double, c_int, dp,
                                    ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
        "zfv_ref_set_threads": (c_int, [vp, c_int]),
        "zfv_ref_eval": (c_int, [vp, c_int, dp, dp]),
        "zfv_ref_geometry": (c_int, [vp, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_uint8)]),
        "zfv_ref_counts": (c_int, [vp, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
        "zfv_ref_get_stencils": (c_int, [vp, c_int, ip, c_int, ip, ip, c_int]),
        "zfv_ref_ldlt_solve": (c_int, [c_int, dp, c_int, dp]),
        "zfv_ref_rank": (c_int, [c_int, c_int, dp]),
        "zfv_ref_hllc": (c_int, [c_double, dp, dp, dp]),
        "zfv_ref_rusanov": (c_int, [c_double, dp, dp, dp]),
        "zfv_ref_rule": (c_int, [c_int, c_int, dp, dp, c_int]),
        "zfv_ref_last_error": (ctypes.c_char_p, []),
        "zfv_ref_kat_error": (ctypes.c_char_p, []),
        "zfv_ref_eq_solve": (c_int, [vp, c_int, dp, c_int, c_double, c_double, dp, dp, ctypes.POINTER(c_int)]),
        "zfv_ref_eq_extrapolate": (c_int, [vp, dp, dp, c_int, dp, dp]),
        "zfv_ref_eq_average": (c_int, [vp, dp, dp, c_int, dp, c_int, dp]),
        "zfv_ref_poly_eval": (c_int, [c_int, c_int, dp, dp, dp, c_double, dp, dp]),
        "zfv_ref_poly_lincomb": (c_int, [c_int, ip, ip, dp, dp, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + (load().zfv_ref_last_error() or b"").decode())


class Oracle:
    """The reference's RateOfChange / SSP3 / LocalCFL restated on the CPU."""

    def __init__(self, mesh, params, threads: int = 1):
        self.lib = load()
        self.mesh = mesh
        self._cm = mesh.c_struct()
        h = ctypes.c_void_p()
        _check(self.lib.zfv_ref_create(ctypes.byref(self._cm), ctypes.byref(params), ctypes.byref(h)))
        self.h = h
        self.lib.zfv_ref_set_threads(h, threads)

    def close(self):
        if self.h:
            self.lib.zfv_ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n: int):
        self.lib.zfv_ref_set_threads(self.h, n)

    def rate(self, u, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        _check(self.lib.zfv_ref_rate(self.h, _dp(u), _dp(out), t))
        return out

    def step(self, u, t, dt):
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        _check(self.lib.zfv_ref_step(self.h, _dp(u), t, dt))
        return u

    def cfl_dt(self, u, cfl):
        u = np.ascontiguousarray(u, dtype=np.float64)
        dt = ctypes.c_double()
        _check(self.lib.zfv_ref_cfl_dt(self.h, _dp(u), cfl, ctypes.byref(dt)))
        return dt.value

    def advance(self, u, cfl, t0, t_end, max_steps=0, drho_crit_rel=0.0, max_exponent=0):
        """TimeLoop with LocalCFL (and RejectLargeDensityChange when
        drho_crit_rel > 0); returns (u, t_final, steps, rejected)."""
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        t = ctypes.c_double()
        k, r = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.zfv_ref_advance(self.h, _dp(u), cfl, t0, t_end, max_steps, drho_crit_rel, max_exponent,
                                        ctypes.byref(t), ctypes.byref(k), ctypes.byref(r)))
        return u, t.value, k.value, r.value

    def set_steady_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._steady = u
        _check(self.lib.zfv_ref_set_steady_state(self.h, _dp(u)))

    def eval(self, i, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(5)
        _check(self.lib.zfv_ref_eval(self.h, i, _dp(x), _dp(out)))
        return out

    def geometry(self):
        n = self.mesh.n_cells
        vol, ctr, inr, ln = np.zeros(n), np.zeros((n, 3)), np.zeros(n), np.zeros(n)
        gh = np.zeros(n, dtype=np.uint8)
        self.lib.zfv_ref_geometry(self.h, _dp(vol), _dp(ctr), _dp(inr), _dp(ln),
                                  gh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return dict(volumes=vol, centers=ctr, inradii=inr, lengths=ln, ghost=gh)

    def counts(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.lib.zfv_ref_counts(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def stencils(self, i):
        l2g = (ctypes.c_int32 * 512)()
        size = ctypes.c_int32()
        mem = (ctypes.c_int32 * 1024)()
        w = self.lib.zfv_ref_get_stencils(self.h, i, l2g, 512, ctypes.byref(size), mem, 1024)
        assert w >= 0
        return list(l2g[: size.value]), list(mem[:w])

    def topology(self):
        n = self.mesh.n_cells
        k = self.mesh.n_dims + 1
        n_int, n_ext = self.counts()
        nbr = np.zeros((n, k), dtype=np.int32)
        eidx = np.zeros((n, k), dtype=np.int32)
        normals = np.zeros((n_int + n_ext, 3))
        ip = ctypes.POINTER(ctypes.c_int32)
        self.lib.zfv_ref_topology.argtypes = [ctypes.c_void_p, ip, ip, ctypes.POINTER(ctypes.c_double)]
        self.lib.zfv_ref_topology(self.h, nbr.ctypes.data_as(ip), eidx.ctypes.data_as(ip), _dp(normals))
        return nbr, eidx, normals


def ldlt_solve(M, b):
    M = np.ascontiguousarray(M, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    n = M.shape[0]
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    _check(load().zfv_ref_ldlt_solve(n, _dp(M), nrhs, _dp(b)))
    return b


def rank(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    return load().zfv_ref_rank(A.shape[0], A.shape[1], _dp(A))


def hllc(gamma, uL, uR):
    f = np.zeros(5)
    load().zfv_ref_hllc(gamma, _dp(np.ascontiguousarray(uL, float)), _dp(np.ascontiguousarray(uR, float)), _dp(f))
    return f


def rusanov(gamma, uL, uR):
    f = np.zeros(5)
    load().zfv_ref_rusanov(gamma, _dp(np.ascontiguousarray(uL, float)), _dp(np.ascontiguousarray(uR, float)), _dp(f))
    return f


def rule(kind, deg):
    w = np.zeros(32)
    x = np.zeros(32 * 4)
    n = load().zfv_ref_rule(kind, deg, _dp(w), _dp(x), 32)
    return w[:n].copy(), x[: 4 * n].reshape(n, 4).copy()


def rk_ode(kind, dt, t_end):
    lib = load()
    lib.zfv_ref_rk_ode.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    out = ctypes.c_double()
    lib.zfv_ref_rk_ode(kind, dt, t_end, ctypes.byref(out))
    return out.value


# ---- pieces for the reference's unit-test known answers -----------------------
_QN_F = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))
_QN_IDF = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_double))


def quasi_newton(f, inv_df, x0, atol, max_iter=20):
    """The oracle's quasi_newton on f: R^n -> R^n (n = 1, 2) with the given
    inverse Jacobian inv_df(x, fx) -> dx; returns (x, status, steps) with
    status 0 converged, 1 not converged (x = x0), 2 not converging."""
    n = len(x0)

    def cf(x, fx):
        r = f(np.array([x[k] for k in range(n)]))
        for k in range(n):
            fx[k] = r[k]

    def cidf(x, fx, dx):
        r = inv_df(np.array([x[k] for k in range(n)]), np.array([fx[k] for k in range(n)]))
        for k in range(n):
            dx[k] = r[k]

    lib = load()
    lib.zfv_ref_quasi_newton.argtypes = [ctypes.c_int, _QN_F, _QN_IDF, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    fcb, icb = _QN_F(cf), _QN_IDF(cidf)
    x = np.zeros(n)
    steps = ctypes.c_int()
    st = lib.zfv_ref_quasi_newton(n, fcb, icb, _dp(np.array(x0, float)), _dp(np.array(atol, float)), max_iter,
                                  _dp(x), ctypes.byref(steps))
    return x, st, steps.value


def _p(params):
    from zisafvm_amd import Params  # struct layout only (include/zfv.h)

    return ctypes.byref(params) if isinstance(params, Params) else params


def eq_solve(params, corners, deg, rho_bar, E_bar):
    """LocalEquilibrium::solve on one simplex (corners: (d+1, 3)); returns
    (theta, x_ref, found)."""
    lib = load()
    c = np.ascontiguousarray(corners, float)
    th, xr, found = np.zeros(2), np.zeros(3), ctypes.c_int()
    rc = lib.zfv_ref_eq_solve(_p(params), c.shape[0] - 1, _dp(c), deg, ctypes.c_double(rho_bar),
                              ctypes.c_double(E_bar), _dp(th), _dp(xr), ctypes.byref(found))
    if rc != 0:
        raise RuntimeError(lib.zfv_ref_kat_error().decode())
    return th, xr, bool(found.value)


def eq_extrapolate(params, theta, x_ref, x):
    x = np.ascontiguousarray(np.atleast_2d(x), float)
    out = np.zeros((x.shape[0], 2))
    load().zfv_ref_eq_extrapolate(_p(params), _dp(np.array(theta, float)), _dp(np.array(x_ref, float)),
                                  x.shape[0], _dp(x), _dp(out))
    return out


def eq_average(params, theta, x_ref, corners, deg):
    c = np.ascontiguousarray(corners, float)
    out = np.zeros(2)
    load().zfv_ref_eq_average(_p(params), _dp(np.array(theta, float)), _dp(np.array(x_ref, float)),
                              c.shape[0] - 1, _dp(c), deg, _dp(out))
    return out


def poly_eval(n_dims, degree, coeffs, moments, x, center=(0.0, 0.0, 0.0), length=1.0):
    """PolyND evaluation; coeffs (dof, nv<=5), moments (dof,)."""
    coeffs = np.atleast_2d(np.asarray(coeffs, float))
    if coeffs.shape[0] == 1 and coeffs.shape[1] > 5:
        coeffs = coeffs.T
    dof = coeffs.shape[0]
    a = np.zeros((dof, 5))
    a[:, : coeffs.shape[1]] = coeffs
    out = np.zeros(5)
    load().zfv_ref_poly_eval(n_dims, degree, _dp(a), _dp(np.asarray(moments, float)),
                             _dp(np.asarray(center, float)), ctypes.c_double(length),
                             _dp(np.asarray(x, float)), _dp(out))
    return out


def poly_lincomb(polys, weights, x):
    """sum_k w_k p_k formed as the WENO sum is formed (metadata of the larger
    degree, ties to the right operand), evaluated at x.  polys: list of
    (n_dims, degree, coeffs (dof, nv), moments (dof,))."""
    maxc = 35
    nd = np.array([p[0] for p in polys], dtype=np.int32)
    dg = np.array([p[1] for p in polys], dtype=np.int32)
    a = np.zeros((len(polys), maxc, 5))
    c = np.zeros((len(polys), maxc))
    for j, (_, _, co, mo) in enumerate(polys):
        co = np.atleast_2d(np.asarray(co, float))
        a[j, : co.shape[0], : co.shape[1]] = co
        c[j, : len(mo)] = mo
    out = np.zeros(5)
    ip = ctypes.POINTER(ctypes.c_int32)
    load().zfv_ref_poly_lincomb(len(polys), nd.ctypes.data_as(ip), dg.ctypes.data_as(ip), _dp(a), _dp(c),
                                _dp(np.asarray(weights, float)), _dp(np.asarray(x, float)), _dp(out))
    return out
