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
