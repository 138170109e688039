"""Asynchronous snapshots (zfv_snapshot_begin / zfv_snapshot_end): the
snapshot holds the state as of the begin call, bit for bit, while the
instance keeps stepping; the steps after it are unaffected; one snapshot at a
time."""
import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import workloads as W

pytestmark = pytest.mark.gpu


def _solver(wl):
    s = z.Solver(wl.mesh, wl.params)
    s.set_state(wl.u0)
    if wl.steady is not None:
        s.set_steady_state(wl.steady)
    return s


@pytest.mark.parametrize("name,make", [("vortex2d", lambda: W.vortex_2d(n=64)),
                                       ("polytrope3d", lambda: W.polytrope_3d(n=12, amplitude=1e-3))])
def test_snapshot_is_the_state_at_begin(gpu, name, make):
    wl = make()
    a, b = _solver(wl), _solver(wl)
    dt = a.cfl_dt(0.45)
    a.steps(3, 0.0, dt)
    b.steps(3, 0.0, dt)
    at3 = b.get_state()
    a.snapshot_begin()
    a.steps(5, 3 * dt, dt)  # the run goes on while the copy drains
    snap = a.snapshot_end()
    np.testing.assert_array_equal(snap, at3)
    b.steps(5, 3 * dt, dt)
    np.testing.assert_array_equal(a.get_state(), b.get_state())
    # a second snapshot reuses the buffers
    a.snapshot_begin()
    np.testing.assert_array_equal(a.snapshot_end(), b.get_state())


def test_one_snapshot_at_a_time(gpu):
    wl = W.vortex_2d(n=16)
    s = _solver(wl)
    with pytest.raises(z.ZfvError):
        s.snapshot_end()
    s.snapshot_begin()
    with pytest.raises(z.ZfvError):
        s.snapshot_begin()
    np.testing.assert_array_equal(s.snapshot_end(), s.get_state())
