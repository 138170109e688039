"""Synthetic workloads of the benchmark configurations: meshes, parameters
and initial conditions (cell averages by the cell quadrature, as the
reference's experiments compute them: ``u0(i) = average(cell, ic)``,
src/zisa/experiments/polytrope.cpp:47-50).

The reference has neither a mesh generator nor the isentropic-vortex or
isothermal-atmosphere initial conditions; those are the workload's own
(textbook) definitions, recorded here with their parameters.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import (
    ZFV_EQ_ISENTROPIC,
    ZFV_EQ_ISOTHERMAL,
    ZFV_FLUX_HLLC,
    ZFV_FLUX_RUSANOV,
    ZFV_GRAVITY_AXIAL,
    ZFV_GRAVITY_POLYTROPE,
    Mesh,
    Params,
    ball_mesh_3d,
    lattice_mesh_2d,
    make_params,
)

# Cell rules (weight, barycentric point): Dunavant degrees 1-5 and Shunn-Ham
# degrees 1-3, the tables of src/zisa/math/{triangular,tetrahedral}_rule.cpp.
def _perm3(w, lam):
    if len(lam) == 1:
        return [(w, (lam[0],) * 3)]
    a, b = lam
    return [(w, (a, b, b)), (w, (b, a, b)), (w, (b, b, a))]


TRI_RULES = {
    1: _perm3(1.0, [1.0 / 3.0]),
    2: _perm3(1.0 / 3.0, [2.0 / 3.0, 1.0 / 6.0]),
    3: _perm3(-0.5625, [1.0 / 3.0]) + _perm3(1.5625 / 3.0, [0.6, 0.2]),
    4: _perm3(0.109951743655322, [0.816847572980459, 0.091576213509771])
    + _perm3(0.223381589678011, [0.108103018168070, 0.445948490915965]),
    5: _perm3(0.225, [1.0 / 3.0]) + _perm3(0.132394152788506, [0.059715871789770, 0.470142064105115])
    + _perm3(0.125939180544827, [0.797426985353087, 0.101286507323456]),
}
_a, _b = 0.5854101966249680, 0.1381966011250110
_c, _d, _e, _f = 0.7784952948213300, 0.0738349017262234, 0.4062443438840510, 0.0937556561159491
TET_RULES = {
    1: [(1.0, (0.25, 0.25, 0.25, 0.25))],
    2: [(0.25, p) for p in ((_a, _b, _b, _b), (_b, _a, _b, _b), (_b, _b, _a, _b), (_b, _b, _b, _a))],
    3: [(0.0476331348432089, p) for p in ((_c, _d, _d, _d), (_d, _c, _d, _d), (_d, _d, _c, _d), (_d, _d, _d, _c))]
    + [(0.1349112434378610, p) for p in ((_e, _e, _f, _f), (_e, _f, _e, _f), (_e, _f, _f, _e),
                                          (_f, _e, _e, _f), (_f, _e, _f, _e), (_f, _f, _e, _e))],
}
_TRI4 = TRI_RULES[4]
_TET3 = TET_RULES[3]


def cell_average(mesh: Mesh, f, deg: Optional[int] = None) -> np.ndarray:
    """Average of f(x) (x: (n,3) -> (n,5)) over every cell with the cell rule
    of degree `deg` (default: 4 in 2D, 3 in 3D — the workloads' volume rule)."""
    if mesh.n_dims == 2:
        rule = TRI_RULES[deg or 4]
    else:
        rule = TET_RULES[deg or 3]
    acc = None
    for w, lam in rule:
        x = sum(l * mesh.corners[:, k, :] for k, l in enumerate(lam))
        v = w * f(x)
        acc = v if acc is None else acc + v
    return acc


def conserved(rho, vx, vy, vz, p, gamma):
    E = p / (gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, E], axis=-1)


def isentropic_vortex(x, beta=5.0, center=(0.5, 0.5), radius=0.1, gamma=1.4, background=(1.0, 1.0, 1.0, 1.0)):
    """Shu's isentropic vortex with core radius `radius` on background
    (rho, u, v, p); exact solution translates with (u, v)."""
    rho0, u0, v0, p0 = background
    dx = (x[:, 0] - center[0]) / radius
    dy = (x[:, 1] - center[1]) / radius
    r2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2 * np.pi) * f * dy
    dv = beta / (2 * np.pi) * f * dx
    dT = -(gamma - 1.0) * beta ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    T = p0 / rho0 + dT
    rho = rho0 * (T * rho0 / p0) ** (1.0 / (gamma - 1.0))
    p = rho * T
    return conserved(rho, u0 + du, v0 + dv, 0.0 * rho, p, gamma)


def isothermal_atmosphere(x, amplitude=0.0, width=0.1, center=(0.5, 0.5), gamma=1.4):
    """rho = p = exp(-y) (RT = 1, gravity phi = y) with a Gaussian pressure bump."""
    y = x[:, 1]
    rho = np.exp(-y)
    r2 = ((x[:, 0] - center[0]) ** 2 + (x[:, 1] - center[1]) ** 2) / width ** 2
    p = np.exp(-y) * (1.0 + amplitude * np.exp(-r2))
    z = 0.0 * y
    return conserved(rho, z, z, z, p, gamma)


def polytrope_n1(x, amplitude=0.0, width=0.2, rhoC=1.0, K=1.0, G=1.0):
    """n = 1 polytrope (gamma = 2): rho = rhoC sin(a r)/(a r), p = K rho^2,
    a = sqrt(2 pi G / K) (include/zisa/model/gravity_impl.hpp:31-59), with a
    Gaussian radial-velocity bump."""
    alpha = np.sqrt(2.0 * np.pi * G / K)
    r = np.sqrt((x ** 2).sum(axis=1))
    ar = alpha * r + 1e-300
    rho = rhoC * np.sin(ar) / ar
    p = K * rho ** 2
    vr = amplitude * np.exp(-(r / width) ** 2)
    rs = np.maximum(r, 1e-300)
    return conserved(rho, vr * x[:, 0] / rs, vr * x[:, 1] / rs, vr * x[:, 2] / rs, p, 2.0)


@dataclass
class Workload:
    name: str
    mesh: Mesh
    params: Params
    u0: np.ndarray
    steady: Optional[np.ndarray]
    t_end: Optional[float] = None
    n_steps: Optional[int] = None
    cfl: float = 0.45
    meta: dict = field(default_factory=dict)


def vortex_2d(n=128, order=3, beta=5.0, seed=1, jitter=0.2, radius=0.1) -> Workload:
    """Plain method: 2D Euler, isentropic vortex, periodic unit square."""
    mesh = lattice_mesh_2d(n, jitter=jitter, seed=seed, periodic=(True, True))
    p = make_params(n_dims=2, order=order, flux=ZFV_FLUX_HLLC)
    u0 = cell_average(mesh, lambda x: isentropic_vortex(x, beta=beta, radius=radius))
    return Workload("vortex2d", mesh, p, u0, None, t_end=0.1, cfl=0.45,
                    meta=dict(n=n, jitter=jitter, seed=seed, beta=beta, core_radius=radius, order=order))


def isothermal_2d(n=360, order=4, amplitude=1e-6, seed=2, jitter=0.2, ghost_layers=3) -> Workload:
    """Well-balanced 2D: isothermal atmosphere rho = p = exp(-y), gravity (0,-1),
    periodic in x, frozen ghost layers in y."""
    mesh = lattice_mesh_2d(n, jitter=jitter, seed=seed, periodic=(True, False), ghost_layers=ghost_layers)
    p = make_params(n_dims=2, order=order, flux=ZFV_FLUX_HLLC, equilibrium=ZFV_EQ_ISOTHERMAL,
                    gravity=ZFV_GRAVITY_AXIAL, gravity_g=1.0, gravity_axis=(0.0, 1.0, 0.0))
    u0 = cell_average(mesh, lambda x: isothermal_atmosphere(x, amplitude=amplitude))
    steady = cell_average(mesh, lambda x: isothermal_atmosphere(x, amplitude=0.0))
    return Workload("isothermal2d", mesh, p, u0, steady, t_end=0.25, cfl=0.45,
                    meta=dict(n=n, jitter=jitter, seed=seed, amplitude=amplitude, order=order))


def isentropic_atmosphere(x, g=1.0, K=1.0, gamma=1.4, h0=3.5, amplitude=0.0, width=0.1, center=(0.5, 0.5)):
    """Isentropic atmosphere under gravity phi = g y: enthalpy h = h0 - g y,
    rho = ((gamma-1)/(gamma K) h)^(1/(gamma-1)), p = K rho^gamma
    (include/zisa/model/isentropic_equilibrium.hpp:44-63,
    include/zisa/model/ideal_gas_eos.hpp:162-168), with a Gaussian pressure bump."""
    y = x[:, 1]
    rho = ((gamma - 1.0) / (gamma * K) * (h0 - g * y)) ** (1.0 / (gamma - 1.0))
    r2 = ((x[:, 0] - center[0]) ** 2 + (x[:, 1] - center[1]) ** 2) / width ** 2
    p = K * rho ** gamma * (1.0 + amplitude * np.exp(-r2))
    zz = 0.0 * y
    return conserved(rho, zz, zz, zz, p, gamma)


def isentropic_2d(n=16, order=3, amplitude=0.0, gamma=1.4, seed=6, jitter=0.2, flux_bc=2, ghost_layers=0) -> Workload:
    """Well-balanced 2D with the reference's isentropic equilibrium for a
    general gamma (its per-point power laws): periodic in x, solid-wall rows at
    y = 0 and y = 1 closed by the equilibrium flux BC (flux_bc 2) unless
    ghost layers are asked for."""
    mesh = lattice_mesh_2d(n, jitter=jitter, seed=seed, periodic=(True, False), ghost_layers=ghost_layers)
    p = make_params(n_dims=2, order=order, flux=ZFV_FLUX_HLLC, equilibrium=ZFV_EQ_ISENTROPIC,
                    gravity=ZFV_GRAVITY_AXIAL, gravity_g=1.0, gravity_axis=(0.0, 1.0, 0.0), gamma=gamma,
                    flux_bc=flux_bc)
    u0 = cell_average(mesh, lambda x: isentropic_atmosphere(x, gamma=gamma, amplitude=amplitude))
    steady = cell_average(mesh, lambda x: isentropic_atmosphere(x, gamma=gamma))
    return Workload("isentropic2d", mesh, p, u0, steady if ghost_layers > 0 else None, t_end=0.1, cfl=0.45,
                    meta=dict(n=n, jitter=jitter, seed=seed, amplitude=amplitude, order=order, gamma=gamma))


def batched_2d(n_instances=64, min_cells=5000, max_cells=200000, seed=4, order=3):
    """Batched workload: instances with triangle counts log-uniform in
    [min_cells, max_cells]; even members are isentropic vortices (strength
    U[2, 6]), odd members isothermal atmospheres with a pressure bump of
    amplitude log-uniform in [1e-7, 1e-3].  All CWENO-AO order 3."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_instances):
        cells = np.exp(rng.uniform(np.log(min_cells), np.log(max_cells)))
        n = max(8, int(round(np.sqrt(cells / 2.0))))
        if k % 2 == 0:
            wl = vortex_2d(n=n, order=order, beta=float(rng.uniform(2.0, 6.0)), seed=1000 + k)
        else:
            amp = float(np.exp(rng.uniform(np.log(1e-7), np.log(1e-3))))
            wl = isothermal_2d(n=n, order=order, amplitude=amp, seed=1000 + k)
        wl.meta["member"] = k
        out.append(wl)
    return out


def serving_requests(n_requests=16, seed=5, min_tris=10000, max_tris=200000, min_tets=10000, max_tets=60000,
                     t_min=0.05, t_max=0.5):
    """Solve requests for the serving runtime: every fourth a 3D polytrope
    (tetrahedra log-uniform in [min_tets, max_tets]), the others 2D, half
    isentropic vortices and half isothermal atmospheres with a pressure bump
    (triangles log-uniform in [min_tris, max_tris]); end times U[t_min, t_max];
    CWENO-AO order 3 throughout."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_requests):
        t_end = float(rng.uniform(t_min, t_max))
        if k % 4 == 3:
            tets = np.exp(rng.uniform(np.log(min_tets), np.log(max_tets)))
            # Kuhn cubes inside the unit ball: about 6 (n/2)^3 (4/3) pi tetrahedra
            n = max(8, int(round(2.0 * (tets / (8.0 * np.pi)) ** (1.0 / 3.0))))
            wl = polytrope_3d(n=n, amplitude=float(10 ** rng.uniform(-7, -3)), seed=2000 + k)
            wl.n_steps = 0
        else:
            tris = np.exp(rng.uniform(np.log(min_tris), np.log(max_tris)))
            n = max(8, int(round(np.sqrt(tris / 2.0))))
            if k % 2 == 0:
                wl = vortex_2d(n=n, beta=float(rng.uniform(2.0, 6.0)), seed=3000 + k)
            else:
                wl = isothermal_2d(n=n, order=3, amplitude=float(10 ** rng.uniform(-7, -3)), seed=3000 + k)
        wl.t_end = t_end
        out.append(wl)
    return out


def polytrope_3d(n=40, order=3, amplitude=1e-5, seed=3, jitter=0.1, radius=1.0, ghost_radius=0.85) -> Workload:
    """Well-balanced 3D: n = 1 polytrope in the unit ball, radial gravity from
    the enclosed mass, Rusanov flux, isentropic equilibrium."""
    mesh = ball_mesh_3d(n, radius=radius, ghost_radius=ghost_radius, jitter=jitter, seed=seed)
    p = make_params(n_dims=3, order=order, flux=ZFV_FLUX_RUSANOV, equilibrium=ZFV_EQ_ISENTROPIC,
                    gravity=ZFV_GRAVITY_POLYTROPE, polytrope=(1.0, 1.0, 1.0), gamma=2.0)
    u0 = cell_average(mesh, lambda x: polytrope_n1(x, amplitude=amplitude))
    steady = cell_average(mesh, lambda x: polytrope_n1(x, amplitude=0.0))
    return Workload("polytrope3d", mesh, p, u0, steady, n_steps=50, cfl=0.45,
                    meta=dict(n=n, jitter=jitter, seed=seed, amplitude=amplitude, order=order))
