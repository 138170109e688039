"""Pins of the CPU oracle's building blocks against the reference's own unit
tests (known-answer tests) and against what the mathematics fixes.

Each test names the oracle function it pins and the reference test or closed
form it uses.
"""
import numpy as np
import pytest

from oracle import zfv_ref as R


def euler_flux(u, gamma):
    rho, m1, m2, m3, E = u
    p = (E - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / rho) * (gamma - 1.0)
    v = m1 / rho
    return np.array([m1, v * m1 + p, v * m2, v * m3, v * (E + p)])


def test_hllc_consistency_reference_kat():
    # test/zisa/unit_test/flux/hllc.cpp:10-26: F_HLLC(u, u) = F(u) to 1e-12,
    # gamma = 1.6, u = (1, -0.2, 0.3, 0.8, 12)
    u = np.array([1.0, -0.2, 0.3, 0.8, 12.0])
    f = R.hllc(1.6, u, u)
    np.testing.assert_allclose(f, euler_flux(u, 1.6), rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", range(5))
def test_hllc_consistency_random(seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.1, 2)
    m = rng.normal(size=3)
    E = 0.5 * (m @ m) / rho + rng.uniform(0.5, 3)
    u = np.array([rho, *m, E])
    np.testing.assert_allclose(R.hllc(1.4, u, u), euler_flux(u, 1.4), rtol=1e-13, atol=1e-13)


def test_hllc_rest_state_is_pressure():
    # equal states at rest: the star correction vanishes (cK = 1), F = (0, p, 0, 0, 0)
    u = np.array([0.7, 0.0, 0.0, 0.0, 2.5])
    p = 2.5 * (1.4 - 1.0)
    np.testing.assert_array_equal(R.hllc(1.4, u, u), [0.0, p, 0.0, 0.0, 0.0])


def test_hllc_mirror_symmetry():
    # reflecting the normal direction: F(uL, uR) = -M F(M uR, M uL), M flips m_n
    rng = np.random.default_rng(7)
    for _ in range(10):
        uL = np.array([rng.uniform(0.5, 2), *rng.normal(size=3), 0.0])
        uR = np.array([rng.uniform(0.5, 2), *rng.normal(size=3), 0.0])
        for u in (uL, uR):
            u[4] = 0.5 * (u[1:4] @ u[1:4]) / u[0] + rng.uniform(0.5, 2)
        M = np.diag([1, -1, 1, 1, 1.0])
        np.testing.assert_allclose(R.hllc(1.4, uL, uR), -M @ R.hllc(1.4, M @ uR, M @ uL), rtol=1e-12, atol=1e-12)


def test_hllc_supersonic_upwinding():
    # sL >= 0: the flux is the left physical flux
    uL = np.array([1.0, 5.0, 0.0, 0.0, 0.5 * 25 + 2.5])
    uR = np.array([0.5, 2.5, 0.0, 0.0, 0.5 * 12.5 + 1.0])
    np.testing.assert_allclose(R.hllc(1.4, uL, uR), euler_flux(uL, 1.4), rtol=1e-14)


def test_rusanov_consistency_and_symmetry():
    u = np.array([1.0, -0.2, 0.3, 0.8, 12.0])
    np.testing.assert_allclose(R.rusanov(1.6, u, u), euler_flux(u, 1.6), atol=1e-13)
    uL = np.array([1.0, 0.3, 0.1, 0.0, 3.0])
    uR = np.array([0.8, -0.1, 0.2, 0.0, 2.0])
    M = np.diag([1, -1, 1, 1, 1.0])
    np.testing.assert_allclose(R.rusanov(1.4, uL, uR), -M @ R.rusanov(1.4, M @ uR, M @ uL), atol=1e-14)


def test_gauss_legendre_nodes_reference_table():
    # test/zisa/unit_test/math/gauss_legendre.cpp:10-57 (4 and 5 point tables
    # to 2e-14); the edge rule maps deg -> deg/2+1 points, weights 0.5*GL
    w, x = R.rule(0, 7)  # 4 points
    ref_x = [-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053]
    ref_w = [0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454]
    np.testing.assert_allclose(x[:, 0], ref_x, atol=2e-14)
    np.testing.assert_allclose(2 * w, ref_w, atol=2e-14)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5, 6, 7])
def test_edge_rule_weights_sum_to_one(deg):
    # test/zisa/unit_test/math/edge_rule.cpp:12-38
    w, _ = R.rule(0, deg)
    assert abs(w.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5])
def test_triangle_rule_exact_for_monomials(deg):
    # exact integration of x^a y^b, a+b <= deg, over the reference triangle:
    # int = a! b! / (a+b+2)!  (area 1/2), rule weights are normalized to 1
    from math import factorial

    w, lam = R.rule(1, deg)
    x, y = lam[:, 1], lam[:, 2]
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2) * 2.0
            assert abs((w * x ** a * y ** b).sum() - exact) < 5e-14, (a, b)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_tetrahedron_rule_exact_for_monomials(deg):
    from math import factorial

    w, lam = R.rule(2, deg)
    x, y, z = lam[:, 1], lam[:, 2], lam[:, 3]
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3) * 6.0
                assert abs((w * x ** a * y ** b * z ** c).sum() - exact) < 1e-14


@pytest.mark.parametrize("n,seed", [(3, 0), (5, 1), (9, 2)])
def test_ldlt_solves_spd_systems(n, seed):
    # Eigen::LDLT restated: A^T A x = b solved to round-off vs numpy
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n + 4, n))
    M = A.T @ A
    b = rng.normal(size=(n, 3))
    x = R.ldlt_solve(M, b)
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=1e-10, atol=1e-12)


def test_ldlt_pivots_on_largest_diagonal():
    # a permutation-sensitive case: tiny leading diagonal
    M = np.array([[1e-12, 1.0, 0.0], [1.0, 4.0, 1.0], [0.0, 1.0, 3.0]])
    b = np.array([[1.0], [2.0], [3.0]])
    x = R.ldlt_solve(M, b)
    np.testing.assert_allclose(M @ x, b, atol=1e-9)


def test_rank():
    A = np.array([[1.0, 2.0], [2.0, 4.0], [3.0, 6.0]])
    assert R.rank(A) == 1
    A[0, 1] = 2.5
    assert R.rank(A) == 2


def test_ssp3_constant_rhs_is_exact():
    # test/zisa/unit_test/ode/runge_kutta.cpp:91-110: constant RHS integrated
    # exactly to 1e-12
    assert abs(R.rk_ode(2, 0.1, 1.0) - 1.0) < 1e-12


@pytest.mark.parametrize("kind,exact", [(0, np.exp(1.0)), (1, 1.0 / 3.0)])
def test_ssp3_observed_order(kind, exact):
    # test/zisa/unit_test/ode/runge_kutta.cpp:172-199: SSP3 order in
    # [2.9, 3.1] with dt 0.1 -> 0.01 on u' = u and u' = t^2
    if kind == 1:
        # SSP3 integrates u' = t^2 exactly (a polynomial of degree 2 in t)
        assert abs(R.rk_ode(1, 0.1, 1.0) - exact) < 1e-14
        return
    e1 = abs(R.rk_ode(kind, 0.1, 1.0) - exact)
    e2 = abs(R.rk_ode(kind, 0.01, 1.0) - exact)
    rate = np.log10(e1 / e2)
    assert 2.9 <= rate <= 3.1, rate


def test_two_triangle_grid_reference_kat():
    # test/zisa/unit_test/grid/grid.cpp:21-107: neighbours, edge numbering
    # (interior first), the diagonal's normal and the volumes
    import zisafvm_amd as z
    from oracle.zfv_ref import Oracle

    verts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0]])
    ids = np.array([[0, 1, 3], [1, 2, 3]], dtype=np.int32)
    mesh = z.Mesh(2, ids, verts[ids], np.zeros(2, dtype=np.uint8))
    p = z.make_params(n_dims=2, order=1, n_biased=0)
    o = Oracle(mesh, p)
    nbr, eidx, normals = o.topology()
    assert nbr.tolist() == [[-1, 1, -1], [-1, -1, 0]]
    assert eidx[0, 1] == 0 and eidx[1, 2] == 0
    assert (eidx[0, 0], eidx[0, 2], eidx[1, 0], eidx[1, 1]) == (1, 2, 3, 4)
    assert normals.shape[0] == 5
    np.testing.assert_allclose(normals[0], [0.5 / np.sqrt(0.5), 0.5 / np.sqrt(0.5), 0.0], atol=1e-12)
    np.testing.assert_allclose(o.geometry()["volumes"], [0.5, 0.5], atol=1e-12)


# ---- quasi_newton (include/zisa/math/quasi_newton.hpp:12-50) -------------------
def test_quasi_newton_scalar_reference_kat():
    # test/zisa/unit_test/math/quasi_newton.cpp:9-23: x^2 - 1.2345^2 from 1.9,
    # exact inverse derivative, atol 1e-10: converges to 10 atol of the root
    root = 1.2345
    x, st, _ = R.quasi_newton(lambda x: x * x - root * root, lambda x, fx: fx / (2.0 * x), [1.9], [1e-10])
    assert st == 0
    assert abs(x[0] - root) < 10.0 * 1e-10


def test_quasi_newton_system_reference_kat():
    # test/zisa/unit_test/math/quasi_newton.cpp:25-53: f(h, K) = (hK + h - K, K)
    # with its exact inverse Jacobian, from (0.2, -0.4): converges to 0
    def f(x):
        h, K = x
        return np.array([h * K + h - K, K])

    def inv_df(x, fx):
        h, K = x
        rho, E = fx
        inv_det = 1.0 / (K + 1.0)
        return np.array([inv_det * (rho + (-h + 1.0) * E), inv_det * ((K + 1.0) * E)])

    x, st, _ = R.quasi_newton(f, inv_df, [0.2, -0.4], [1e-10, 1e-10])
    assert st == 0
    assert np.linalg.norm(x) < 10.0 * 1e-10


def test_quasi_newton_without_root_gives_back_the_guess():
    # no real root (x^2 + 1): Newton wanders; the solve reports failure and
    # returns the initial guess, as the reference does
    x, st, steps = R.quasi_newton(lambda x: x * x + 1.0, lambda x, fx: fx / (2.0 * x), [0.7], [1e-10])
    assert st != 0 and x[0] == 0.7 and steps >= 5


def test_quasi_newton_stops_at_max_iter():
    # linear convergence (a root of multiplicity 3 with the simple-root step)
    # does not reach 1000 atol in 20 steps: not converged, the guess returned
    x, st, steps = R.quasi_newton(lambda x: x ** 3, lambda x, fx: fx / (3.0 * x * x), [1.0], [1e-13])
    assert st == 1 and steps == 20 and x[0] == 1.0


# ---- LocalEquilibrium (include/zisa/model/local_equilibrium_impl.hpp:41-93) ----
def _isentropic_radial():
    import zisafvm_amd as z

    # IdealGasEOS(gamma 1.2, R 0.9), ConstantGravityRadial(0.9):
    # test/zisa/unit_test/model/local_equilibrium.cpp:15-22
    p = z.make_params(n_dims=2, equilibrium=z.ZFV_EQ_ISENTROPIC, gravity=z.ZFV_GRAVITY_RADIAL, gravity_g=0.9,
                      gamma=1.2, gas_constant=0.9)
    return p


def test_local_equilibrium_recovery_reference_kat():
    # test/zisa/unit_test/model/local_equilibrium.cpp:15-60: the cell average
    # of the equilibrium theta_ref = (h, K) = (10, 3) anchored at (0.5, 0.6)
    # over a small triangle at (1, 1), solved back from that average alone,
    # extrapolates to a far point and averages over another triangle to 1e-9
    p = _isentropic_radial()
    theta_ref, x_ref = np.array([10.0, 3.0]), np.array([0.5, 0.6, 0.0])
    tri_ref = np.array([[1.0, 1.0, 0.0], [1.01, 1.0, 0.0], [1.0, 1.01, 0.0]])
    rho_bar, E_bar = R.eq_average(p, theta_ref, x_ref, tri_ref, 2)
    theta, xr, found = R.eq_solve(p, tri_ref, 2, rho_bar, E_bar)
    assert found
    xy = np.array([1.1, 2.1, 0.0])
    approx = R.eq_extrapolate(p, theta, xr, xy)[0]
    exact = R.eq_extrapolate(p, theta_ref, x_ref, xy)[0]
    assert np.linalg.norm(approx - exact) < 1e-9
    tri = np.array([[1.2, 1.1, 0.0], [1.21, 1.1, 0.0], [1.2, 1.11, 0.0]])
    approx = R.eq_average(p, theta, xr, tri, 2)
    exact = R.eq_average(p, theta_ref, x_ref, tri, 2)
    assert np.linalg.norm(approx - exact) < 1e-9


def test_isentropic_extrapolation_is_isentropic_and_hydrostatic():
    # what the physics fixes (include/zisa/model/isentropic_equilibrium.hpp:44-63,
    # ideal_gas_eos.hpp:162-168): along the equilibrium p / rho^gamma = K and
    # the enthalpy h = gamma/(gamma-1) p/rho satisfies h + phi = const; a
    # wrong exponent or sign in the extrapolation breaks one of them
    p = _isentropic_radial()
    gamma, g = 1.2, 0.9
    theta, x_ref = np.array([10.0, 3.0]), np.array([0.5, 0.6, 0.0])
    rng = np.random.default_rng(3)
    x = np.c_[rng.uniform(-2, 2, (20, 2)), np.zeros(20)]
    rhoE = R.eq_extrapolate(p, theta, x_ref, x)
    rho, pres = rhoE[:, 0], rhoE[:, 1] * (gamma - 1.0)
    np.testing.assert_allclose(pres / rho ** gamma, 3.0, rtol=1e-12)
    h = gamma / (gamma - 1.0) * pres / rho
    np.testing.assert_allclose(h + g * np.linalg.norm(x, axis=1), 10.0 + g * np.linalg.norm(x_ref), rtol=1e-12)


def test_isothermal_extrapolation_closed_form():
    # the isothermal equilibrium (an extension): rho = rho_ref exp(-(phi - phi_ref)/RT),
    # E = rho RT / (gamma - 1)
    import zisafvm_amd as z

    p = z.make_params(n_dims=2, equilibrium=z.ZFV_EQ_ISOTHERMAL, gravity=z.ZFV_GRAVITY_AXIAL, gravity_g=2.0,
                      gravity_axis=(0.0, 1.0, 0.0), gamma=1.4)
    theta, x_ref = np.array([1.3, 0.7]), np.array([0.2, 0.1, 0.0])
    x = np.array([[0.0, 0.5, 0.0], [1.0, -0.3, 0.0]])
    rhoE = R.eq_extrapolate(p, theta, x_ref, x)
    rho = 1.3 * np.exp(-(2.0 * x[:, 1] - 2.0 * 0.1) / 0.7)
    np.testing.assert_allclose(rhoE[:, 0], rho, rtol=1e-14)
    np.testing.assert_allclose(rhoE[:, 1], rho * 0.7 / 0.4, rtol=1e-14)


# ---- PolyND (include/zisa/math/poly2d_impl.hpp:271-323) ------------------------
def test_poly2d_examples_reference_kat():
    # test/zisa/unit_test/math/poly2d.cpp:94-145: coefficients 1..6, moments
    # (0, 0, 0, 1, 2, 3), centre 0, length 1
    a = np.arange(1.0, 7.0)
    c = np.array([0.0, 0.0, 0.0, 1.0, 2.0, 3.0])
    assert abs(R.poly_eval(2, 2, a[:, None], c, [0.0, 0.0, 0.0])[0] - (1.0 - 4.0 - 10.0 - 18.0)) < 1e-14
    x, y = -3.0, 2.0
    exact = 1.0 + 2.0 * x + 3.0 * y + 4.0 * (x * x - 1.0) + 5.0 * (x * y - 2.0) + 6.0 * (y * y - 3.0)
    assert abs(R.poly_eval(2, 2, a[:, None], c, [x, y, 0.0])[0] - exact) < 1e-14
    # saxpy-like: 0.2 p + q - 0.4 p, q = 1 + 2x + 3y (degree 1, zero moments)
    xq = np.array([-3.4, 2.138, 0.0])
    p = (2, 2, a[:, None], c)
    q = (2, 1, np.array([[1.0], [2.0], [3.0]]), np.zeros(3))
    approx = R.poly_lincomb([p, q, p], [0.2, 1.0, -0.4], xq)[0]
    exact = 0.2 * R.poly_eval(2, 2, a[:, None], c, xq)[0] + R.poly_eval(2, 1, q[2], q[3], xq)[0] \
        - 0.4 * R.poly_eval(2, 2, a[:, None], c, xq)[0]
    assert abs(approx - exact) < 1e-14


def test_poly3d_examples_reference_kat():
    # test/zisa/unit_test/math/poly2d.cpp:193-283: two variables, degree 2,
    # moments (0, 0, 0, 0, 1, ..., 6), at (-3, 2, 0.2)
    coeffs = np.array([[1.0, -1.0], [2.0, -1.0], [-1.0, -1.0], [-3.0, -1.0], [2.0, -3.0], [4.0, -2.0],
                       [2.0, -5.0], [1.0, -3.0], [1.0, -3.0], [3.0, -1.0]])
    mom = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    x, y, z = -3.0, 2.0, 0.2
    basis = np.array([1.0, x, y, z, x * x - 1.0, x * y - 2.0, x * z - 3.0, y * y - 4.0, y * z - 5.0, z * z - 6.0])
    exact = basis @ coeffs
    approx = R.poly_eval(3, 2, coeffs, mom, [x, y, z])[:2]
    assert np.linalg.norm(approx - exact) < 1e-14
    # saxpy-like with q (degree 1, zero moments)
    qc = np.array([[1.0, -1.0], [-1.0, -1.0], [2.0, -1.0], [3.0, -1.0]])
    xq = np.array([-3.4, 2.138, 0.2])
    p3 = (3, 2, coeffs, mom)
    q3 = (3, 1, qc, np.zeros(4))
    approx = R.poly_lincomb([p3, q3, p3], [0.2, 1.0, -0.4], xq)[:2]
    pe = R.poly_eval(3, 2, coeffs, mom, xq)[:2]
    qe = R.poly_eval(3, 1, qc, np.zeros(4), xq)[:2]
    assert np.linalg.norm(approx - (0.2 * pe + qe - 0.4 * pe)) < 1e-14


def test_kernel_poly_eval_matches_reference_examples():
    # the kernels' own evaluation (poly_eval in zisafvm_amd/csrc/stage_math.hpp,
    # run by the host emulation) on the same examples
    import ctypes

    import emu_binding as E

    lib = E.load()
    dp = ctypes.POINTER(ctypes.c_double)
    lib.zfv_emu_poly_eval.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]

    def ev(nd, deg, a, c, X):
        out = ctypes.c_double()
        a, c, X = (np.ascontiguousarray(v, float) for v in (a, c, X))
        assert lib.zfv_emu_poly_eval(nd, deg, a.ctypes.data_as(dp), c.ctypes.data_as(dp), X.ctypes.data_as(dp),
                                     ctypes.byref(out)) == 0
        return out.value

    a = np.arange(1.0, 7.0)
    c = np.array([0.0, 0.0, 0.0, 1.0, 2.0, 3.0])
    x, y = -3.0, 2.0
    exact = 1.0 + 2.0 * x + 3.0 * y + 4.0 * (x * x - 1.0) + 5.0 * (x * y - 2.0) + 6.0 * (y * y - 3.0)
    assert abs(ev(2, 2, a, c, [x, y, 0.0]) - exact) < 1e-14
    coeffs = np.array([1.0, 2.0, -1.0, -3.0, 2.0, 4.0, 2.0, 1.0, 1.0, 3.0])
    mom = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    x, y, z = -3.0, 2.0, 0.2
    basis = np.array([1.0, x, y, z, x * x - 1.0, x * y - 2.0, x * z - 3.0, y * y - 4.0, y * z - 5.0, z * z - 6.0])
    assert abs(ev(3, 2, coeffs, mom, [x, y, z]) - basis @ coeffs) < 1e-14


# ---- central stencils (src/zisa/reconstruction/stencil.cpp:159-399) -------------
@pytest.mark.parametrize("n_dims,order", [(2, 3), (2, 4), (3, 3)])
def test_central_stencil_holds_every_nearer_cell(n_dims, order):
    # test/zisa/unit_test/reconstruction/stencil.cpp:10-41: every cell whose
    # centre is closer to the cell's centre than the central stencil's
    # farthest member belongs to the central stencil (for every cell with a
    # full-order stencil here, not only cell 42)
    import zisafvm_amd as z
    from oracle.zfv_ref import Oracle

    if n_dims == 2:
        mesh = z.lattice_mesh_2d(20, jitter=0.2, seed=21, periodic=(False, False), ghost_layers=3)
    else:
        mesh = z.ball_mesh_3d(10, radius=1.0, ghost_radius=0.7, jitter=0.1, seed=21)
    o = Oracle(mesh, z.make_params(n_dims=n_dims, order=order))
    ctr = o.geometry()["centers"]
    checked = 0
    for i in range(mesh.n_cells):
        _, mem = o.stencils(i)
        if mem[0] != order:
            continue  # first-order ghost stencils
        ids = mem[2 : 2 + mem[1]]
        d = np.linalg.norm(ctr - ctr[i], axis=1)
        nearer = set(np.nonzero(d < d[ids].max())[0].tolist())
        assert nearer <= set(ids), (i, sorted(nearer - set(ids)))
        checked += 1
    assert checked > 100
