"""CPU oracle for the hot path of MongeAmpere++ -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
(``mongeampere_amd``) never imports it.

* ``kantorovich`` / ``first_moment`` call ``liboracle.so`` (``ma_oracle.c``):
  an FP64 restatement of ``include/MA/kantorovich.hpp:35-141`` and
  ``include/MA/lloyd.hpp:30-69`` with exact Side1/2/3 predicates
  (``include/MA/predicates.hpp:21-168``).
* ``solve_laplacian_matrix`` restates ``include/MA/optimal_transport.hpp:41-87``
  (gauge fix: drop the last row/column, sparse direct solve, d[N-1] = 0).
* ``ot_solve`` restates ``include/MA/optimal_transport.hpp:89-193`` (damped
  Newton with the eps0 = min(m0, nu)/2 floor and the (1 - alpha/2) decrease
  test), plus an optional max-relative stopping rule and an alpha floor.
"""
from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        lib.orc_run.restype = P
        lib.orc_run.argtypes = [ctypes.c_int, I64, P, P, I64, P, I64, P, P]
        lib.orc_status.restype = ctypes.c_int
        lib.orc_status.argtypes = [P]
        lib.orc_ntriplets.restype = I64
        lib.orc_ntriplets.argtypes = [P]
        lib.orc_copy.restype = None
        lib.orc_copy.argtypes = [P, P, P, P, P, P, P, P]
        lib.orc_free.restype = None
        lib.orc_free.argtypes = [P]
        lib.orc_copy_mom2.restype = None
        lib.orc_copy_mom2.argtypes = [P, P]
        lib.orc_nfrag.restype = I64
        lib.orc_nfrag.argtypes = [P, P]
        lib.orc_copy_frags.restype = None
        lib.orc_copy_frags.argtypes = [P, P, P, P, P]
        lib.orc_draw_laguerre.restype = ctypes.c_int
        lib.orc_draw_laguerre.argtypes = ([I64, P, P, I64, P, I64, P, P, P, ctypes.c_int]
                                          + [ctypes.c_double] * 4 + [ctypes.c_int] * 2 + [P])
        for nm in ("orc_side1", "orc_side1_exact"):
            f = getattr(lib, nm)
            f.restype = ctypes.c_int
            f.argtypes = [P, P, ctypes.c_double, ctypes.c_double]
        for nm in ("orc_side2", "orc_side2_exact"):
            f = getattr(lib, nm)
            f.restype = ctypes.c_int
            f.argtypes = [P, P, P, P]
        for nm in ("orc_side3", "orc_side3_exact"):
            f = getattr(lib, nm)
            f.restype = ctypes.c_int
            f.argtypes = [P, P] + [ctypes.c_double] * 4 + [P]
        lib.orc_ac6_poly.restype = ctypes.c_double
        lib.orc_ac6_poly.argtypes = [P, P]
        lib.orc_centroid_lin.restype = ctypes.c_double
        lib.orc_centroid_lin.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _LIB = lib
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


@dataclass
class OracleStats:
    pairs: int = 0
    fragments: int = 0
    predicates: int = 0
    exact: int = 0
    cells: int = 0
    neighbours: int = 0
    seconds: float = 0.0


def _run(mode, mesh, X, w):
    lib = _lib()
    vxy = _c(mesh.vxy, np.float64)
    vval = _c(mesh.vval, np.float64)
    tri = _c(mesh.tri, np.int32)
    X = _c(X, np.float64).reshape(-1, 2)
    w = _c(w, np.float64).reshape(-1)
    n = X.shape[0]
    if w.shape[0] != n:
        raise ValueError("weights and Diracs differ in length")
    t0 = time.perf_counter()
    r = lib.orc_run(mode, vxy.shape[0], _ptr(vxy), _ptr(vval), tri.shape[0], _ptr(tri), n,
                    _ptr(X), _ptr(w))
    dt = time.perf_counter() - t0
    if not r:
        raise MemoryError("oracle allocation failed")
    try:
        st = lib.orc_status(r)
        if st != 0:
            raise RuntimeError(f"oracle failed with status {st}")
        nt = lib.orc_ntriplets(r)
        fval = np.zeros(1)
        g = np.zeros(n)
        mom = np.zeros((n, 2))
        ti = np.zeros(max(nt, 1), np.int32)
        tj = np.zeros(max(nt, 1), np.int32)
        tv = np.zeros(max(nt, 1))
        stats = np.zeros(6, np.int64)
        lib.orc_copy(r, _ptr(fval), _ptr(g), _ptr(mom), _ptr(ti), _ptr(tj), _ptr(tv), _ptr(stats))
        mom2 = np.zeros((n, 3))
        if mode == 2:
            lib.orc_copy_mom2(r, _ptr(mom2))
    finally:
        lib.orc_free(r)
    s = OracleStats(*[int(v) for v in stats], seconds=dt)
    if mode == 2:
        return float(fval[0]), g, (mom, mom2), (ti[:nt], tj[:nt], tv[:nt]), s
    return float(fval[0]), g, mom, (ti[:nt], tj[:nt], tv[:nt]), s


def kantorovich(mesh, X, w, return_stats=False):
    """Value, gradient (cell masses) and Hessian of the Kantorovich functional
    (``include/MA/kantorovich.hpp:35-141``).  The Hessian is assembled from the
    triplets like ``setFromTriplets`` (duplicates summed, ``:137-139``)."""
    fval, g, _, (ti, tj, tv), st = _run(0, mesh, X, w)
    n = g.shape[0]
    H = sp.csr_matrix((tv, (ti, tj)), shape=(n, n))
    H.sum_duplicates()
    H.sort_indices()
    return (fval, g, H, st) if return_stats else (fval, g, H)


def kantorovich_triplets(mesh, X, w):
    """Raw triplets (row, col, value) in emission order."""
    fval, g, _, trip, st = _run(0, mesh, X, w)
    return fval, g, trip, st


def first_moment(mesh, X, w):
    """Cell masses and first moments (``include/MA/lloyd.hpp:30-69``)."""
    _, m, mom, _, _ = _run(1, mesh, X, w)
    return m, mom


def second_moment(mesh, X, w):
    """Cell masses, first moments and second moments (int rho x^2, int rho y^2,
    int rho x y), ``include/MA/lloyd.hpp:71-123``."""
    _, m, (mom, mom2), _, _ = _run(2, mesh, X, w)
    return m, mom, mom2


def fragments(mesh, X, w):
    """Polygon of every non-empty fragment (Laguerre cell of a Dirac cut by a
    triangle): the traversal ``voronoi_triangulation_intersection``
    (``include/MA/voronoi_triangulation_intersection.hpp:315-343``) with a
    callback that records ``(polygon, face, vertex)``.  Returns
    ``(cell, tri, start, xy)``: fragment f is Dirac ``cell[f]`` cut by
    triangle ``tri[f]``, vertices ``xy[start[f]:start[f+1]]`` counter-clockwise,
    in the order the oracle visits them (Dirac by Dirac)."""
    lib = _lib()
    vxy = _c(mesh.vxy, np.float64)
    vval = _c(mesh.vval, np.float64)
    tri = _c(mesh.tri, np.int32)
    X = _c(X, np.float64).reshape(-1, 2)
    w = _c(w, np.float64).reshape(-1)
    n = X.shape[0]
    r = lib.orc_run(3, vxy.shape[0], _ptr(vxy), _ptr(vval), tri.shape[0], _ptr(tri), n,
                    _ptr(X), _ptr(w))
    if not r:
        raise MemoryError("oracle allocation failed")
    try:
        st = lib.orc_status(r)
        if st != 0:
            raise RuntimeError(f"oracle failed with status {st}")
        nvx = np.zeros(1, np.int64)
        nf = lib.orc_nfrag(r, _ptr(nvx))
        cell = np.zeros(max(nf, 1), np.int32)
        ft = np.zeros(max(nf, 1), np.int32)
        start = np.zeros(nf + 1, np.int64)
        xy = np.zeros((max(int(nvx[0]), 1), 2))
        lib.orc_copy_frags(r, _ptr(cell), _ptr(ft), _ptr(start), _ptr(xy))
    finally:
        lib.orc_free(r)
    return cell[:nf], ft[:nf], start, xy[:int(nvx[0])]


def draw_laguerre_diagram(mesh, X, w, colors, box, width, height):
    """``draw_laguerre_diagram`` (``include/MA/rasterization.hpp:491-547``)
    with a summing ``put_pixel``: image ``(height, width, channels)`` where
    pixel (i, j) (row j counted from ``box``'s lower edge) receives, from
    every fragment, its exact coverage of the pixel times the mean of the
    density over the fragment's vertices times ``colors[cell]``."""
    lib = _lib()
    vxy = _c(mesh.vxy, np.float64)
    vval = _c(mesh.vval, np.float64)
    tri = _c(mesh.tri, np.int32)
    X = _c(X, np.float64).reshape(-1, 2)
    w = _c(w, np.float64).reshape(-1)
    col = np.asarray(colors, np.float64)
    if col.ndim == 1:
        col = col[:, None]
    col = np.ascontiguousarray(col)
    nc = col.shape[1]
    img = np.zeros((height, width, nc))
    st = lib.orc_draw_laguerre(vxy.shape[0], _ptr(vxy), _ptr(vval), tri.shape[0], _ptr(tri),
                               X.shape[0], _ptr(X), _ptr(w), _ptr(col), nc, float(box[0]),
                               float(box[1]), float(box[2]), float(box[3]), int(width),
                               int(height), _ptr(img))
    if st != 0:
        raise RuntimeError(f"oracle failed with status {st}")
    return img


def voronoi_polygon_intersection(P, X, w):
    """Laguerre cell of every Dirac clipped to the polygon P, in exact
    rational arithmetic (small inputs): ``voronoi_polygon_intersection``
    (``include/MA/voronoi_polygon_intersection.hpp:153-188``) clips P by the
    power half-plane of every RT neighbour of v with Sutherland-Hodgman
    (``polygon_halfplane_intersection``, ``include/MA/polygon_intersection.hpp:46-73``);
    clipping by all other Diracs (a superset of the neighbours) gives the same
    polygon.  Half-plane of j: 2 x.(y_j - y_i) <= |y_j|^2 - |y_i|^2 - w_j + w_i.
    Repeated and collinear vertices are removed from the result.  Returns a
    list of (k, 2) float arrays, counter-clockwise."""
    from fractions import Fraction as Fr

    Pq = [(Fr(float(x)), Fr(float(y))) for x, y in np.asarray(P, np.float64).reshape(-1, 2)]
    a2 = sum(Pq[i][0] * Pq[(i + 1) % len(Pq)][1] - Pq[(i + 1) % len(Pq)][0] * Pq[i][1]
             for i in range(len(Pq)))
    if a2 < 0:
        Pq = Pq[::-1]
    Y = [(Fr(float(x)), Fr(float(y))) for x, y in np.asarray(X, np.float64).reshape(-1, 2)]
    W = [Fr(float(v)) for v in np.asarray(w, np.float64).reshape(-1)]
    out = []
    for i in range(len(Y)):
        poly = list(Pq)
        for j in range(len(Y)):
            if j == i or not poly:
                continue
            ax, ay = 2 * (Y[j][0] - Y[i][0]), 2 * (Y[j][1] - Y[i][1])
            c = (Y[j][0] ** 2 + Y[j][1] ** 2) - (Y[i][0] ** 2 + Y[i][1] ** 2) - W[j] + W[i]
            f = [ax * p[0] + ay * p[1] - c for p in poly]  # <= 0 inside
            nxt = []
            for k in range(len(poly)):
                k1 = (k + 1) % len(poly)
                if f[k] <= 0:
                    nxt.append(poly[k])
                if (f[k] < 0 < f[k1]) or (f[k1] < 0 < f[k]):
                    t = f[k] / (f[k] - f[k1])
                    nxt.append((poly[k][0] + t * (poly[k1][0] - poly[k][0]),
                                poly[k][1] + t * (poly[k1][1] - poly[k][1])))
            poly = nxt
        # drop repeated and collinear vertices
        changed = True
        while changed and len(poly) >= 3:
            changed = False
            for k in range(len(poly)):
                a, b, c2 = poly[k - 1], poly[k], poly[(k + 1) % len(poly)]
                cr = (b[0] - a[0]) * (c2[1] - b[1]) - (b[1] - a[1]) * (c2[0] - b[0])
                if b == a or cr == 0:
                    del poly[k]
                    changed = True
                    break
        if len(poly) < 3:
            poly = []
        out.append(np.array([[float(x), float(y)] for x, y in poly]).reshape(-1, 2))
    return out


def lloyd(mesh, X, w):
    """Cell masses and centroids (``include/MA/lloyd.hpp:126-144``)."""
    m, mom = first_moment(mesh, X, w)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = mom / m[:, None]
    return m, c


def side1(v, w, p):
    return _lib().orc_side1(_ptr(_c(v, np.float64)), _ptr(_c(w, np.float64)), p[0], p[1])


def side2(v, u2, u3, w):
    a = [_c(x, np.float64) for x in (v, u2, u3, w)]
    return _lib().orc_side2(*[_ptr(x) for x in a])


def side3(v, u, p, q, w):
    a = [_c(x, np.float64) for x in (v, u, w)]
    return _lib().orc_side3(_ptr(a[0]), _ptr(a[1]), p[0], p[1], q[0], q[1], _ptr(a[2]))


def side1_exact(v, w, p):
    return _lib().orc_side1_exact(_ptr(_c(v, np.float64)), _ptr(_c(w, np.float64)), p[0], p[1])


def side2_exact(v, u2, u3, w):
    a = [_c(x, np.float64) for x in (v, u2, u3, w)]
    return _lib().orc_side2_exact(*[_ptr(x) for x in a])


def side3_exact(v, u, p, q, w):
    a = [_c(x, np.float64) for x in (v, u, w)]
    return _lib().orc_side3_exact(_ptr(a[0]), _ptr(a[1]), p[0], p[1], q[0], q[1], _ptr(a[2]))


def ac6_poly(tri6, c10):
    return _lib().orc_ac6_poly(_ptr(_c(tri6, np.float64)), _ptr(_c(c10, np.float64)))


def centroid_lin(tri6, a, b, c):
    return _lib().orc_centroid_lin(_ptr(_c(tri6, np.float64)), a, b, c)


# --------------------------------------------------------------------------
# Newton driver (include/MA/optimal_transport.hpp)
# --------------------------------------------------------------------------

def solve_laplacian_matrix(h, g, verbose=False):
    """``include/MA/optimal_transport.hpp:41-87``: solve the gauge-fixed system
    (last row and column dropped) with a sparse direct factorisation; the last
    entry of the result is 0."""
    n = h.shape[0]
    hs = sp.csc_matrix(h[: n - 1, : n - 1])
    gs = np.asarray(g[: n - 1], dtype=np.float64)
    lu = spla.splu(hs, permc_spec="COLAMD")
    ds = lu.solve(gs)
    err = float(np.linalg.norm(hs @ ds - gs))
    d = np.zeros(n)
    d[: n - 1] = ds
    return d, err


@dataclass
class SolveStats:
    niter: int = 0
    neval: int = 0
    status: str = "ok"
    history: list = field(default_factory=list)
    seconds: float = 0.0
    eval_seconds: float = 0.0
    solve_seconds: float = 0.0


def ot_solve(mesh, X, nu, x=None, eps_g=1e-7, maxiter=100, verbose=False, stop="l2",
             tol=None, alpha_min=2.0 ** -30):
    """``include/MA/optimal_transport.hpp:89-193``.

    stop="l2": the reference rule, continue while ||m - nu||_2 >= eps_g.
    stop="maxrel": continue while max_i |m_i - nu_i| / nu_i > tol.
    Returns (weights, SolveStats).  An empty initial cell (eps0 <= 0,
    ``:137-148``) returns status "empty_cell" with the weights untouched."""
    nu = np.asarray(nu, dtype=np.float64)
    n = X.shape[0]
    st = SolveStats()
    t_start = time.perf_counter()

    def f(xv):
        st.neval += 1
        t0 = time.perf_counter()
        fval, g, h = kantorovich(mesh, X, xv)
        st.eval_seconds += time.perf_counter() - t0
        m = g.copy()
        return fval - nu.dot(xv), m, g - nu, h

    if x is None or len(x) != n:
        x = np.zeros(n)
    x = np.array(x, dtype=np.float64)
    fx, m, g, h = f(x)
    eps0 = min(m.min(), nu.min()) / 2.0
    if eps0 <= 0:
        st.status = "empty_cell"
        st.seconds = time.perf_counter() - t_start
        return x, st

    def keep_going():
        if stop == "l2":
            return np.linalg.norm(g) >= eps_g
        return np.max(np.abs(g) / nu) > tol

    while keep_going():
        st.niter += 1
        if st.niter - 1 > maxiter:  # `niter++ <= maxiter` (:151)
            st.status = "not_converged"
            break
        t0 = time.perf_counter()
        d, _ = solve_laplacian_matrix(h, g)
        st.solve_seconds += time.perf_counter() - t0
        d = -d
        alpha = 1.0
        x0 = x.copy()
        n0 = np.linalg.norm(g)
        while True:
            x = x0 + alpha * d
            fx, m, g, h = f(x)
            if m.min() >= eps0 and np.linalg.norm(g) <= (1 - alpha / 2) * n0:
                break
            alpha *= 0.5
            if alpha < alpha_min:
                st.status = "linesearch_fail"
                break
        st.history.append((st.niter, fx, float(np.linalg.norm(g)), float(np.max(np.abs(g) / nu)),
                           alpha, st.neval))
        if verbose:
            print(f"it {st.niter}: f={fx} |df|={np.linalg.norm(g)} tau={alpha} eval={st.neval}")
        if st.status != "ok":
            break
    st.seconds = time.perf_counter() - t_start
    return x, st
