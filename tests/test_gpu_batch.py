"""Batched execution: every member of a batch ends bit-identical to the same
instance advanced alone by its own time loop (CFL each step, clamped to
t_end), and members of different sizes / physics coexist."""
import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import workloads as W

pytestmark = pytest.mark.gpu


def alone(wl, t_end, max_steps):
    s = z.Solver(wl.mesh, wl.params)
    s.set_state(wl.u0)
    if wl.steady is not None:
        s.set_steady_state(wl.steady)
    t, k = 0.0, 0
    while t < t_end and k < max_steps:
        dt = min(s.cfl_dt(wl.cfl), t_end - t)
        s.step(t, dt)
        t += dt
        k += 1
    return s.get_state(), t, k


@pytest.mark.parametrize("xcd", [True, False])
def test_batch_members_equal_their_own_time_loops(gpu, xcd):
    wls = [W.vortex_2d(n=12, beta=4.0), W.isothermal_2d(n=10, amplitude=1e-3, order=3), W.vortex_2d(n=20, beta=2.0),
           W.isothermal_2d(n=16, amplitude=1e-4, order=3)]
    t_ends = [0.01, 0.02, 0.015, 0.005]
    cap = sum(w.mesh.n_cells for w in wls)
    b = z.Batch(wls[0].params, cell_capacity=cap, face_capacity=2 * cap, max_instances=8)
    b.set_xcd_chunks(xcd)
    solvers, slots = [], []
    for w, te in zip(wls, t_ends):
        s = z.Solver(w.mesh, w.params)
        s.set_state(w.u0)
        if w.steady is not None:
            s.set_steady_state(w.steady)
        solvers.append(s)
        slots.append(b.add(s, w.cfl, 0.0, te, 6))
    b.run(10)
    for w, te, s, slot in zip(wls, t_ends, solvers, slots):
        ref, t_ref, k_ref = alone(w, te, 6)
        q = b.query(slot)
        assert not q["active"]
        assert q["steps"] == k_ref and abs(q["t"] - t_ref) <= 1e-15
        np.testing.assert_array_equal(s.get_state(), ref)


def test_batch_admits_and_retires_between_steps(gpu):
    a, c = W.vortex_2d(n=12), W.vortex_2d(n=16)
    b = z.Batch(a.params, cell_capacity=a.mesh.n_cells + c.mesh.n_cells, face_capacity=4 * c.mesh.n_cells,
                max_instances=2)
    sa = z.Solver(a.mesh, a.params)
    sa.set_state(a.u0)
    slot_a = b.add(sa, 0.45, 0.0, 1.0, 3)
    b.run(2)
    sc = z.Solver(c.mesh, c.params)
    sc.set_state(c.u0)
    slot_c = b.add(sc, 0.45, 0.0, 1.0, 2)
    b.run(5)
    assert b.query(slot_a)["steps"] == 3 and b.query(slot_c)["steps"] == 2
    ra, _, _ = alone(a, 1.0, 3)
    rc, _, _ = alone(c, 1.0, 2)
    np.testing.assert_array_equal(sa.get_state(), ra)
    np.testing.assert_array_equal(sc.get_state(), rc)
    b.remove(slot_a)
    # capacity checks: a third member does not fit in two slots
    sd = z.Solver(a.mesh, a.params)
    sd.set_state(a.u0)
    b.add(sd, 0.45, 0.0, 1.0, 1)
    with pytest.raises(z.ZfvError):
        b.add(z.Solver(a.mesh, a.params), 0.45, 0.0, 1.0, 1)


def test_member_steady_state_snapshot_and_destroy_inside_a_batch(gpu):
    # zfv_set_steady_state on a member reaches the batch (its frozen BC is
    # on: the member still equals its own loop, which has the BC from the
    # start); a snapshot begun before zfv_batch_run holds the state before
    # the run; destroying a member removes it from the batch, and the others
    # keep running
    a, c = W.isothermal_2d(n=10, amplitude=1e-3, order=3), W.vortex_2d(n=12)
    cap = a.mesh.n_cells + c.mesh.n_cells
    b = z.Batch(a.params, cell_capacity=cap, face_capacity=2 * cap, max_instances=4)
    sa, sc = z.Solver(a.mesh, a.params), z.Solver(c.mesh, c.params)
    sa.set_state(a.u0)
    sc.set_state(c.u0)
    slot_a = b.add(sa, a.cfl, 0.0, 1.0, 4)
    slot_c = b.add(sc, c.cfl, 0.0, 1.0, 4)
    sa.set_steady_state(a.steady)
    before = sa.get_state()
    sa.snapshot_begin()
    b.run(2)
    np.testing.assert_array_equal(sa.snapshot_end(), before)
    assert b.query(slot_c)["steps"] == 2
    sc.close()
    with pytest.raises(z.ZfvError):
        b.query(slot_c)
    b.run(5)
    q = b.query(slot_a)
    assert q["steps"] == 4 and not q["active"]
    ref, t_ref, _ = alone(a, 1.0, 4)
    assert abs(q["t"] - t_ref) <= 1e-15
    np.testing.assert_array_equal(sa.get_state(), ref)
    with pytest.raises(z.ZfvError):
        b.add(sa, a.cfl, 0.0, 1.0, 1)  # already a member
    b.close()
    # the batch is gone: its former member can join another one
    b2 = z.Batch(a.params, cell_capacity=cap, face_capacity=2 * cap, max_instances=4)
    b2.add(sa, a.cfl, 0.0, 1.0, 1)
    b2.close()


def test_batches_of_two_shapes_in_flight_together(gpu):
    # a 2D and a 3D batch launched one after the other and waited for
    # together (their streams overlap on the device): every member ends
    # bit-identical to its own time loop; calls on a batch or a member with
    # steps in flight wait for them first
    w2 = [W.vortex_2d(n=12, beta=4.0), W.isothermal_2d(n=10, amplitude=1e-3, order=3)]
    w3 = [W.polytrope_3d(n=10, amplitude=1e-3)]
    batches, members = [], []
    for wls in (w2, w3):
        cap = sum(w.mesh.n_cells for w in wls)
        b = z.Batch(wls[0].params, cell_capacity=cap, face_capacity=4 * cap, max_instances=4)
        for w in wls:
            s = z.Solver(w.mesh, w.params)
            s.set_state(w.u0)
            if w.steady is not None:
                s.set_steady_state(w.steady)
            members.append((w, s, b, b.add(s, w.cfl, 0.0, 1.0, 4)))
        batches.append(b)
    for b in batches:
        b.launch(2)
    for b in batches:
        b.launch(2)
    # a query with steps in flight waits for them
    w, s, b, slot = members[0]
    assert b.query(slot)["steps"] == 4
    for b in batches:
        b.wait()
    for w, s, b, slot in members:
        ref, t_ref, k_ref = alone(w, 1.0, 4)
        q = b.query(slot)
        assert q["steps"] == k_ref == 4 and abs(q["t"] - t_ref) <= 1e-15
        np.testing.assert_array_equal(s.get_state(), ref)
    # a member's state read with its batch in flight waits for the batch
    w, s, b, slot = members[-1]
    b.launch(1)
    u = s.get_state()
    b.wait()
    np.testing.assert_array_equal(u, s.get_state())
    for b in batches:
        b.close()
