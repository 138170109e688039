"""Serving: requests that arrive while others run are admitted as memory
frees up, share the device, and each result is bit-identical to the same
request run alone by its own time loop."""
import time

import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import serving as S
from zisafvm_amd import workloads as W

pytestmark = pytest.mark.gpu


def alone(wl):
    s = z.Solver(wl.mesh, wl.params)
    s.set_state(wl.u0)
    if wl.steady is not None:
        s.set_steady_state(wl.steady)
    t, k = 0.0, 0
    max_steps = wl.n_steps or 1 << 30
    while t < wl.t_end and k < max_steps:
        dt = min(s.cfl_dt(wl.cfl), wl.t_end - t)
        s.step(t, dt)
        t += dt
        k += 1
    return s.get_state(), k


def mix():
    reqs = []
    for k, n in enumerate([10, 14, 8, 12, 16, 9]):
        wl = W.vortex_2d(n=n, beta=3.0 + 0.3 * k) if k % 2 == 0 else W.isothermal_2d(n=n, amplitude=1e-4)
        wl.t_end = 0.004 + 0.002 * k
        reqs.append(wl)
    # a 3D member (its own batch shape)
    wl3 = W.polytrope_3d(n=10, amplitude=1e-3)
    wl3.t_end, wl3.n_steps = 1.0, 3
    reqs.append(wl3)
    return reqs


def test_results_equal_alone_runs(gpu):
    reqs = mix()
    # a budget that holds about two requests at a time forces queueing
    est = [S.estimate_device_bytes(w.mesh, w.params) for w in reqs]
    srv = S.Server(hbm_budget_bytes=2 * max(est), batch_cell_capacity=1 << 16, max_members=8, setup_workers=2)
    rids = []
    for w in reqs:
        rids.append(srv.submit(w.mesh, w.params, w.u0, w.steady, w.t_end, w.cfl, w.n_steps or 0))
        time.sleep(0.01)
    srv.wait(rids, timeout_s=120)
    for w, r in zip(reqs, rids):
        st = srv.poll(r)
        assert st.state == S.DONE
        ref, k = alone(w)
        assert st.steps == k
        np.testing.assert_array_equal(srv.result(r, w.mesh.n_cells), ref)
    stats = srv.stats()
    assert stats.reserved_bytes == 0 and stats.running == 0
    srv.close()


def test_oversized_request_is_rejected(gpu):
    wl = W.vortex_2d(n=8)
    srv = S.Server(hbm_budget_bytes=1 << 16, batch_cell_capacity=1 << 12, max_members=2, setup_workers=1)
    with pytest.raises(z.ZfvError):
        srv.submit(wl.mesh, wl.params, wl.u0, None, 0.01)
    srv.close()
