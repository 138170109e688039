"""Writes vortex_convergence.json: the oracle's L1 density errors for the
translating isentropic vortex over a mesh sequence, for CWENO-AO orders 3
and 4 with SSP3 time stepping.

The exact solution is the initial condition translated by (t, t) on the
periodic unit square, so the error of each run is measured against the
cell averages of the shifted initial condition.  The observed orders are
the log2 ratios of successive errors.  The fixture holds the time steps the
runs took, so a GPU run can take the same ones and be compared with the
oracle's errors and orders.

    python tests/golden/make_convergence.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import zisafvm_amd as z  # noqa: E402
from zisafvm_amd import workloads as W  # noqa: E402
from oracle.zfv_ref import Oracle  # noqa: E402

T_END = 0.05
MESHES = (32, 64, 128)
# order 4 runs at a smaller CFL number so that the SSP3 time error stays
# below the fourth-order space error on this sequence
CASES = {3: 0.45, 4: 0.15}


def setup(n, order):
    mesh = z.lattice_mesh_2d(n, jitter=0.2, seed=1, periodic=(True, True))
    p = z.make_params(n_dims=2, order=order, flux=z.ZFV_FLUX_HLLC)
    u0 = W.cell_average(mesh, lambda x: W.isentropic_vortex(x, beta=5.0))
    exact = W.cell_average(mesh, lambda x: W.isentropic_vortex((x - T_END) % 1.0, beta=5.0))
    return mesh, p, u0, exact


def l1_rho(mesh, u, exact):
    return float(np.sum(mesh.volumes() * np.abs(u[:, 0] - exact[:, 0])))


def time_steps(dt0):
    out, t = [], 0.0
    while t < T_END - 1e-15:
        dt = min(dt0, T_END - t)
        out.append(dt)
        t += dt
    return out


def main():
    res = {"t_end": T_END, "beta": 5.0, "jitter": 0.2, "seed": 1, "cases": []}
    threads = os.cpu_count() or 1
    for order, cfl in CASES.items():
        errs, runs = [], []
        for n in MESHES:
            mesh, p, u0, exact = setup(n, order)
            o = Oracle(mesh, p, threads=threads)
            dts = time_steps(o.cfl_dt(u0, cfl))
            u, t = u0.copy(), 0.0
            for dt in dts:
                u = o.step(u, t, dt)
                t += dt
            e = l1_rho(mesh, u, exact)
            errs.append(e)
            runs.append({"n": n, "dt": dts[0], "dt_last": dts[-1], "n_steps": len(dts), "l1_rho": repr(e)})
            print("order", order, "n", n, "steps", len(dts), "L1", e, flush=True)
        orders = [float(np.log2(a / b)) for a, b in zip(errs, errs[1:])]
        print("order", order, "observed", orders, flush=True)
        res["cases"].append({"order": order, "cfl": cfl, "runs": runs, "observed_orders": orders})
    with open(os.path.join(HERE, "vortex_convergence.json"), "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
