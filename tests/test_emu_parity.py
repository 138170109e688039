"""CPU-side parity of the product's numerics: the product's own host setup
(zisafvm_amd/csrc/setup.cpp) and the per-cell/per-face functions of the HIP
kernels (zisafvm_amd/csrc/stage_math.hpp), run in host loops by the
test-only emulation, against the oracle.  The GPU tests repeat these checks
through libzfv_hip.so on the device."""
import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import workloads as W
from oracle.zfv_ref import Oracle
from emu_binding import Emu
from conftest import rel_err, wb_err

TOL = 1e-12


def _stencils_equal(o, e, n):
    for i in range(n):
        assert o.stencils(i) == e.stencils(i), i


@pytest.mark.parametrize("n", [12, 20])
def test_vortex(n):
    wl = W.vortex_2d(n=n)
    o, e = Oracle(wl.mesh, wl.params), Emu(wl.mesh, wl.params)
    _stencils_equal(o, e, wl.mesh.n_cells)
    errs = rel_err(e.rate(wl.u0), o.rate(wl.u0))
    assert max(errs) <= TOL, errs


def test_weno_ao_and_order4():
    wl = W.vortex_2d(n=12)
    for p in (z.make_params(n_dims=2, order=3, reconstruction=z.ZFV_WENO_AO), z.make_params(n_dims=2, order=4)):
        o, e = Oracle(wl.mesh, p), Emu(wl.mesh, p)
        errs = rel_err(e.rate(wl.u0), o.rate(wl.u0))
        assert max(errs) <= TOL, errs


def test_rusanov_flux_bc_ghosts():
    # non-periodic mesh with ghost layers and the constant-extrapolation FluxBC
    mesh = z.lattice_mesh_2d(12, jitter=0.2, seed=9, periodic=(False, False), ghost_layers=3)
    u = W.cell_average(mesh, lambda x: W.isentropic_vortex(x, beta=3.0))
    p = z.make_params(n_dims=2, order=3, flux=z.ZFV_FLUX_RUSANOV, flux_bc=1)
    o, e = Oracle(mesh, p), Emu(mesh, p)
    _stencils_equal(o, e, mesh.n_cells)
    errs = rel_err(e.rate(u), o.rate(u), mask=mesh.ghost == 0)
    assert max(errs) <= TOL, errs


@pytest.mark.parametrize("amp", [1e-3, 1e-6, 0.0])
def test_isothermal_well_balanced(amp):
    wl = W.isothermal_2d(n=14, amplitude=amp)
    o, e = Oracle(wl.mesh, wl.params), Emu(wl.mesh, wl.params)
    Lo, Le = o.rate(wl.u0), e.rate(wl.u0)
    errs = wb_err(Le, Lo, wl.mesh, wl.u0, wl.params.gamma)
    assert max(errs) <= TOL, errs


def test_polytrope_3d():
    wl = W.polytrope_3d(n=12, amplitude=1e-3)
    o, e = Oracle(wl.mesh, wl.params), Emu(wl.mesh, wl.params)
    _stencils_equal(o, e, wl.mesh.n_cells)
    errs = wb_err(e.rate(wl.u0), o.rate(wl.u0), wl.mesh, wl.u0, wl.params.gamma)
    assert max(errs) <= TOL, errs
