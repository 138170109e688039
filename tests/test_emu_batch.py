"""The batch kernels' per-block functions (zisafvm_amd/csrc/batch.hpp), run
on the CPU over the same block tables the library builds: every member of a
batch, each stepping with its own CFL dt clamped to its end time, ends where
the oracle's TimeLoop of that member alone ends, for every reconstruction
block size and with and without the XCD-chunked table."""
import numpy as np
import pytest

from zisafvm_amd import workloads as W
from oracle.zfv_ref import Oracle
from emu_binding import Emu, EmuBatch
from conftest import rel_err, wb_err


def _members():
    wls = [W.vortex_2d(n=12, beta=4.0), W.isothermal_2d(n=10, amplitude=1e-3, order=3), W.vortex_2d(n=20, beta=2.0),
           W.isothermal_2d(n=16, amplitude=1e-4, order=3)]
    return wls, [0.01, 0.02, 0.015, 0.005], [6, 6, 6, 6]


def _run(waves, xcd):
    wls, t_ends, ms = _members()
    emus = [Emu(w.mesh, w.params) for w in wls]
    b = EmuBatch(emus, wls, [w.cfl for w in wls], t_ends, ms)
    b.run(8, waves=waves, xcd_chunks=xcd)
    return wls, t_ends, ms, [b.member(k) for k in range(len(wls))], b.counts()


def test_batch_members_match_the_oracle_time_loops():
    wls, t_ends, ms, out, counts = _run(1, True)
    assert all(c > 0 for c in counts)
    for w, te, m, (u, t, steps, active) in zip(wls, t_ends, ms, out):
        o = Oracle(w.mesh, w.params)
        if w.steady is not None:
            o.set_steady_state(w.steady)
        uo, to, ko, _ = o.advance(w.u0, w.cfl, 0.0, te, max_steps=m)
        assert not active and steps == ko and abs(t - to) <= 1e-14 * to, (steps, ko, t, to)
        # the change over the run per unit time, against the oracle's; well-
        # balanced members relative to the pressure terms that cancel
        if w.params.gravity == 0:
            errs = rel_err(u - w.u0, uo - w.u0, mask=w.mesh.ghost == 0)
        else:
            errs = wb_err((u - w.u0) / t, (uo - w.u0) / to, w.mesh, w.u0, w.params.gamma)
        assert max(errs) <= 1e-11, errs


@pytest.mark.parametrize("waves,xcd", [(2, True), (4, True), (1, False), (4, False)])
def test_batch_results_do_not_depend_on_the_launch_geometry(waves, xcd):
    ref = _run(1, True)[3]
    got = _run(waves, xcd)[3]
    for (ua, ta, ka, _), (ub, tb, kb, _) in zip(ref, got):
        assert ta == tb and ka == kb
        np.testing.assert_array_equal(ua, ub)
