"""Writes the golden fixtures in this directory from the CPU oracle.

These are regression pins of the oracle (and a fixed target for the GPU
path): small meshes, their states and the oracle's rate of change and SSP3
step, stored with 17 significant digits.  The reference ships no golden
vectors of its own (its tests load grid files that are not in its tree), so
these vectors come from the oracle, which is pinned separately by the
reference's known-answer tests and by closed forms (tests/test_oracle_*.py).

    python tests/golden/make_golden.py
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


def params_dict(p):
    out = {}
    for name, _ in p._fields_:
        v = getattr(p, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def cases():
    yield "vortex2d_n6", W.vortex_2d(n=6)
    yield "isothermal2d_n6", W.isothermal_2d(n=6, amplitude=1e-3, ghost_layers=3)
    # the jagged ball surface gives skewed, ill-conditioned boundary stencils
    # (kappa(A^T A) ~ 1e5); as in the workloads, those cells are ghosts
    mesh = z.ball_mesh_3d(8, radius=1.0, ghost_radius=0.6, jitter=0.1, seed=5)
    def pulse(x):
        r2 = ((x - np.array([0.1, -0.1, 0.05])) ** 2).sum(axis=1)
        g = np.exp(-r2 / 0.1)
        return W.conserved(1.0 + 0.3 * g, 0.6 + 0.2 * g, -0.4 + 0.1 * g, 0.5 - 0.3 * g, 1.0 + 0.5 * g, 1.4)

    u = W.cell_average(mesh, pulse)
    p = z.make_params(n_dims=3, order=3, flux=z.ZFV_FLUX_RUSANOV, flux_bc=1)
    yield "ball3d_n8_rusanov", W.Workload("ball3d", mesh, p, u, None)


def main():
    for name, wl in cases():
        o = Oracle(wl.mesh, wl.params)
        if wl.steady is not None:
            o.set_steady_state(wl.steady)
        L = o.rate(wl.u0)
        dt = 0.25 * o.cfl_dt(wl.u0, 0.45)
        u1 = o.step(wl.u0, 0.0, dt)
        data = {
            "name": name,
            "n_dims": wl.mesh.n_dims,
            "vertex_ids": wl.mesh.vertex_ids.tolist(),
            "corners": wl.mesh.corners.tolist(),
            "ghost": wl.mesh.ghost.tolist(),
            "params": params_dict(wl.params),
            "u0": wl.u0.tolist(),
            "steady": None if wl.steady is None else wl.steady.tolist(),
            "rate": L.tolist(),
            "dt": dt,
            "step": u1.tolist(),
        }
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(data, f, separators=(",", ":"), default=float)
        print("wrote", name, wl.mesh.n_cells, "cells")


if __name__ == "__main__":
    main()
