"""zfv_advance_noalloc, the allocation-free variant of the device time loop:
through the C-ABI, two calls in a row on one GPU (the second on a smaller
instance, so the shared scratch array holds the first call's contents when
the second starts) give exactly what zfv_advance gives on fresh instances:
bit-identical states, the same final time, the same accepted and rejected
step counts."""
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


def _density_change_of_one_step(wl):
    s = _solver(wl)
    u0 = s.get_state()
    s.step(0.0, s.cfl_dt(wl.cfl))
    return float(np.max(np.abs(s.get_state()[:, 0] - u0[:, 0]) / u0[:, 0]))


def _pair(wl, t_end, max_steps, crit):
    ref = _solver(wl)
    t_ref, k_ref = ref.advance(wl.cfl, 0.0, t_end, max_steps=max_steps, check_every=8, drho_crit_rel=crit,
                               max_exponent=3)
    s = _solver(wl)
    t, k = s.advance_noalloc(wl.cfl, 0.0, t_end, max_steps=max_steps, check_every=8, drho_crit_rel=crit,
                             max_exponent=3)
    print(wl.name, wl.mesh.n_cells, "cells: steps", k, "rejected", s.rejected_steps, "t", t,
          "| zfv_advance", k_ref, ref.rejected_steps, t_ref, flush=True)
    assert k == k_ref and t == t_ref and s.rejected_steps == ref.rejected_steps
    np.testing.assert_array_equal(s.get_state(), ref.get_state())
    return k, s.rejected_steps


def test_noalloc_time_loop_equals_zfv_advance_twice(gpu):
    # first call: a larger well-balanced instance with step rejection on (the
    # state backup lives in the scratch array)
    big = W.isothermal_2d(n=40, amplitude=1e-3, order=3)
    crit = 0.7 * _density_change_of_one_step(big)
    k1, r1 = _pair(big, 0.01, 1000, crit)
    assert r1 > 0
    # second call: a smaller instance, rejection off, while the array still
    # holds the first call's loop state and backup
    small = W.vortex_2d(n=16)
    k2, r2 = _pair(small, 0.02, 1000, 0.0)
    assert k2 > 0 and r2 == 0
    # and rejection on again on the small one
    crit2 = 0.7 * _density_change_of_one_step(small)
    _pair(small, 0.02, 1000, crit2)


def test_noalloc_rejects_a_backup_larger_than_the_scratch(gpu):
    # 2D: 4 doubles per cell; 16 MiB - 256 B holds 524,280 cells
    wl = W.vortex_2d(n=520)  # 540,800 cells
    s = _solver(wl)
    with pytest.raises(z.ZfvError, match="MiB of scratch"):
        s.advance_noalloc(wl.cfl, 0.0, 1.0, max_steps=1, drho_crit_rel=0.1, max_exponent=2)
    # without step rejection the size is not limited
    t, k = s.advance_noalloc(wl.cfl, 0.0, 1.0, max_steps=1)
    assert k == 1
