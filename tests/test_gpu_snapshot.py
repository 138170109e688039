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


def test_restart_from_an_hdf5_snapshot_continues_the_run(gpu, tmp_path):
    # the reference's dump / restart cycle (DumpSnapshot -> LoadSnapshot,
    # include/zisa/io/dump_snapshot_impl.hpp:22-32, src/zisa/io/load_snapshot.cpp:10-15):
    # the state copied off the device behind the steps, written in the
    # AllVariables layout, read back into a fresh instance and continued,
    # ends bit-identical to the uninterrupted run
    from zisafvm_amd import io as zio

    wl = W.isothermal_2d(n=24, amplitude=1e-3, order=3)
    a = z.Solver(wl.mesh, wl.params)
    a.set_state(wl.u0)
    a.set_steady_state(wl.steady)
    dt = a.cfl_dt(wl.cfl)
    a.steps(5, 0.0, dt)
    a.snapshot_begin()
    a.steps(3, 5 * dt, dt)  # the run goes on while the snapshot drains
    snap = a.snapshot_end()
    p = str(tmp_path / "snap.h5")
    zio.write_snapshot(p, snap, 5 * dt, 5)
    zio.write_steady_state(str(tmp_path / "steady.h5"), wl.steady)
    u, t, k = zio.read_snapshot(p)
    assert t == 5 * dt and k == 5
    b = z.Solver(wl.mesh, wl.params)
    b.set_state(u)
    b.set_steady_state(zio.read_snapshot(str(tmp_path / "steady.h5"), with_time=False))
    b.steps(3, t, dt)
    np.testing.assert_array_equal(b.get_state(), a.get_state())
