"""Size-independent properties of the GPU rate of change at the benchmark
sizes (the parity tests against the oracle run on small meshes):

- a uniform state has a zero rate of change (to round-off of the face sums);
- on a periodic mesh the rate is conservative: sum_i vol_i L_i = 0 for every
  conserved variable (interior fluxes telescope, as in the reference's
  flux loop, include/zisa/fvm_loops/flux_loop.hpp:45-104);
- translating the mesh leaves the rate unchanged;
- the rate of a state does not depend on what the instance ran before.
"""
import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import workloads as W
from conftest import pressure_scale

pytestmark = pytest.mark.gpu


def _vortex(n):
    return W.vortex_2d(n=n)


def test_uniform_state_has_zero_rate(gpu):
    wl = _vortex(128)
    s = z.Solver(wl.mesh, wl.params)
    u = np.tile(np.array([1.0, 0.7, -0.3, 0.0, 2.5]), (wl.mesh.n_cells, 1))
    L = s.rate(u)
    # each cell sums F(u) n |f| / vol over its faces: the face normals cancel
    # to round-off of terms of the size of the pressure scale
    scale = pressure_scale(wl.mesh, u, wl.params.gamma)
    print("max |L| for a uniform state", np.abs(L).max(), "scale", scale)
    assert np.abs(L).max() <= 1e-14 * 5 * scale


def test_rate_is_conservative_on_a_periodic_mesh(gpu):
    wl = _vortex(128)
    s = z.Solver(wl.mesh, wl.params)
    L = s.rate(wl.u0)
    vol = wl.mesh.volumes()
    for v in (0, 1, 2, 4):  # rho, m_x, m_y, E (m_z is zero in 2D)
        total = np.sum(vol * L[:, v])
        size = np.sum(vol * np.abs(L[:, v]))
        print("variable", v, "sum vol L", total, "relative", total / size)
        assert abs(total) <= 1e-13 * size


def test_translation_leaves_the_rate_unchanged(gpu):
    wl = _vortex(64)
    s = z.Solver(wl.mesh, wl.params)
    L0 = s.rate(wl.u0)
    shifted = z.Mesh(wl.mesh.n_dims, wl.mesh.vertex_ids, wl.mesh.corners + np.array([0.25, -0.5, 0.0])[: wl.mesh.corners.shape[-1]],
                     wl.mesh.ghost)
    s2 = z.Solver(shifted, wl.params)
    L1 = s2.rate(wl.u0)
    cols = [0, 1, 2, 4]  # m_z is zero in 2D
    err = np.abs(L1 - L0)[:, cols].max(axis=0) / np.abs(L0)[:, cols].max(axis=0)
    print("translation: rel change per variable", err)
    assert err.max() <= 1e-10


def test_rate_is_repeatable_after_steps(gpu):
    wl = W.isothermal_2d(n=360, amplitude=1e-6)
    s = z.Solver(wl.mesh, wl.params)
    s.set_steady_state(wl.steady)
    L = s.rate(wl.u0)
    s.set_state(wl.u0)
    s.steps(3, 0.0, s.cfl_dt(0.45))
    np.testing.assert_array_equal(s.rate(wl.u0), L)
