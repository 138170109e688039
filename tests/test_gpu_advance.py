"""The device time loop (zfv_advance): LocalCFL steps chosen on the device,
clamped to the end time, must give exactly what the host loop gives with
zfv_cfl_dt + zfv_step (src/zisa/fvm_loops/time_loop.cpp:36-81): the same
number of steps, the same final time and bit-identical states; with step
rejection on, also the same number of rejected steps."""
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


def _host_loop_with_rejection(wl, cfl, t_end, max_steps, drho_crit_rel, max_exponent):
    """TimeLoop with RejectLargeDensityChange on the host
    (src/zisa/fvm_loops/time_loop.cpp:36-81 and 182-196,
    src/zisa/ode/step_rejection.cpp:17-57), each step through zfv_step."""
    s = _solver(wl)
    factor, good, min_factor = 1.0, True, 2.0 ** (-max_exponent)
    t, k, rejected = 0.0, 0, 0
    dt = min(factor * s.cfl_dt(cfl), t_end - t)
    while t < t_end and k < max_steps:
        u0 = s.get_state()
        s.step(t, dt)
        u1 = s.get_state()
        ok = not np.any(np.abs(u0[:, 0] - u1[:, 0]) > drho_crit_rel * u0[:, 0])
        if not ok:
            good = False
            factor /= 2.0
        else:
            if good:
                factor *= 2.0
            good = True
        factor = max(min_factor, min(factor, 1.0))
        if not ok:
            rejected += 1
            s.set_state(u0)
            dt = min(factor * dt, t_end - t)
            continue
        t += dt
        k += 1
        dt = min(factor * s.cfl_dt(cfl), t_end - t)
    return s.get_state(), t, k, rejected


def _first_step_density_change(wl):
    s = _solver(wl)
    u0 = s.get_state()
    s.step(0.0, s.cfl_dt(wl.cfl))
    return float(np.max(np.abs(s.get_state()[:, 0] - u0[:, 0]) / u0[:, 0]))


@pytest.mark.parametrize(
    "name,make,t_end",
    [
        ("vortex2d", lambda: W.vortex_2d(n=32), 0.02),
        ("isothermal2d", lambda: W.isothermal_2d(n=24, amplitude=1e-3, order=3), 0.01),
        ("polytrope3d", lambda: W.polytrope_3d(n=10, amplitude=1e-3), 0.05),
    ],
)
@pytest.mark.parametrize("check_every", [1, 16])
def test_device_step_rejection_equals_host_loop(gpu, name, make, t_end, check_every):
    wl = make()
    # a criterion between the change of a full CFL step and of a half step:
    # full steps are rejected, halved ones accepted, and after two good steps
    # the factor returns to 1, so rejections recur through the run
    crit = 0.7 * _first_step_density_change(wl)
    ref, t_ref, k_ref, r_ref = _host_loop_with_rejection(wl, wl.cfl, t_end, 1000, crit, 3)
    s = _solver(wl)
    t, k = s.advance(wl.cfl, 0.0, t_end, max_steps=1000, check_every=check_every, drho_crit_rel=crit,
                     max_exponent=3)
    print(name, "crit", crit, "steps", k, "rejected", s.rejected_steps, "t", t, "host", k_ref, r_ref, t_ref)
    assert r_ref > 0
    assert k == k_ref and t == t_ref and s.rejected_steps == r_ref
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
