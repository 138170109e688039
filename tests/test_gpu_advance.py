"""The device time loop (zfv_advance): LocalCFL steps chosen on the device,
clamped to the end time, must give exactly what the host loop gives with
zfv_cfl_dt + zfv_step (src/zisa/fvm_loops/time_loop.cpp:36-81): the same
number of steps, the same final time and bit-identical states."""
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


def _host_loop(wl, cfl, t_end, max_steps):
    s = _solver(wl)
    t, k = 0.0, 0
    while t < t_end and k < max_steps:
        dt = min(s.cfl_dt(cfl), t_end - t)
        s.step(t, dt)
        t += dt
        k += 1
    return s.get_state(), t, k


@pytest.mark.parametrize(
    "name,make,t_end,max_steps",
    [
        ("vortex2d", lambda: W.vortex_2d(n=32), 0.02, 1000),
        ("isothermal2d", lambda: W.isothermal_2d(n=24, amplitude=1e-3, order=3), 0.01, 1000),
        ("polytrope3d", lambda: W.polytrope_3d(n=10, amplitude=1e-3), 0.05, 7),
    ],
)
@pytest.mark.parametrize("check_every", [1, 16])
def test_device_time_loop_equals_host_loop(gpu, name, make, t_end, max_steps, check_every):
    wl = make()
    ref, t_ref, k_ref = _host_loop(wl, wl.cfl, t_end, max_steps)
    s = _solver(wl)
    t, k = s.advance(wl.cfl, 0.0, t_end, max_steps=max_steps, check_every=check_every)
    print(name, "steps", k, "t", t, "host", k_ref, t_ref)
    assert k == k_ref and t == t_ref
    np.testing.assert_array_equal(s.get_state(), ref)


def test_device_time_loop_continues_and_stops(gpu):
    wl = W.vortex_2d(n=16)
    ref, t_ref, k_ref = _host_loop(wl, wl.cfl, 0.03, 1000)
    s = _solver(wl)
    t1, k1 = s.advance(wl.cfl, 0.0, 0.03, max_steps=5)
    assert k1 == 5
    t2, k2 = s.advance(wl.cfl, t1, 0.03)
    assert k1 + k2 == k_ref and t2 == t_ref
    np.testing.assert_array_equal(s.get_state(), ref)
    # nothing to do past the end time
    assert s.advance(wl.cfl, t2, t2) == (t2, 0)
