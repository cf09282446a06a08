"""Generate exact golden vectors for the Kantorovich functional on tiny inputs.

Independent brute-force method (no neighbour search, no floating point): every
triangle is clipped by the half-planes {pi_i < pi_j} of ALL other Diracs in
exact rational arithmetic (``fractions.Fraction``), with the reference's strict
tie rule (a point with pi_i == pi_j counts as outside,
``include/MA/predicates.hpp:85-86``).  Masses, the value and the first moments
(and the second moments int rho x^2, y^2, xy of include/MA/lloyd.hpp:71-123) are exact
rationals (the centroid rule and the Albrecht-Collatz rule are exact
for P1 / cubic integrands, ``include/MA/quadrature.hpp:23-42,69-77``); the
Hessian entries -rho(mid) |e| / (2 |y_i - y_j|) (``include/MA/kantorovich.hpp:
109-122``) involve square roots and are evaluated with mpmath at 50 digits.

Run:  python tests/golden/make_golden.py   (writes tests/golden/*.json)
"""
from __future__ import annotations

import json
import os
from fractions import Fraction as Fr

import mpmath
import numpy as np

mpmath.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def grid(n, x0=0.0, y0=0.0, x1=1.0, y1=1.0):
    xs = [x0 + (x1 - x0) * i / (n - 1) for i in range(n)]
    ys = [y0 + (y1 - y0) * j / (n - 1) for j in range(n)]
    vxy = [[xs[i], ys[j]] for i in range(n) for j in range(n)]
    tri = []
    for i in range(n - 1):
        for j in range(n - 1):
            a = i * n + j
            tri.append([a, a + n, a + n + 1])
            tri.append([a, a + n + 1, a + 1])
    return vxy, tri


def power(P, y, w):
    return (P[0] - y[0]) ** 2 + (P[1] - y[1]) ** 2 - w


def clip(poly, labs, Y, W, i, j):
    """Keep points with pi_i < pi_j (strict).  poly: list of (x,y) Fractions
    (ccw), labs[k]: label of edge k -> k+1 (j >= 0 Dirac, < 0 triangle)."""
    s = [power(p, Y[j], W[j]) - power(p, Y[i], W[i]) for p in poly]
    n = len(poly)
    if all(v > 0 for v in s):
        return poly, labs
    out, ol = [], []
    for k in range(n):
        k1 = (k + 1) % n
        a, b = poly[k], poly[k1]
        if s[k] > 0:
            out.append(a)
            ol.append(labs[k])
            if not s[k1] > 0:
                t = s[k] / (s[k] - s[k1])
                out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
                ol.append(j)
        elif s[k1] > 0:
            t = s[k] / (s[k] - s[k1])
            out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
            ol.append(labs[k])
    return out, ol


def area(a, b, c):
    return ((b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1])) / 2


def ac6(a, b, c, f):
    u = (b[0] - a[0], b[1] - a[1])
    v = (c[0] - a[0], c[1] - a[1])
    h, s6, t3 = Fr(1, 2), Fr(1, 6), Fr(2, 3)

    def P(cu, cv):
        return (a[0] + cu * u[0] + cv * v[0], a[1] + cu * u[1] + cv * v[1])

    r = (f(P(h, h)) + f(P(h, 0)) + f(P(0, h))) * Fr(1, 30) + \
        (f(P(s6, t3)) + f(P(t3, s6)) + f(P(s6, s6))) * Fr(9, 30)
    return area(a, b, c) * r


def solve(vxy, vval, tri, X, w):
    V = [(Fr(x), Fr(y)) for x, y in vxy]
    F = [Fr(v) for v in vval]
    Y = [(Fr(x), Fr(y)) for x, y in X]
    W = [Fr(v) for v in w]
    n = len(Y)
    mass = [Fr(0)] * n
    intg = [Fr(0)] * n
    mx = [Fr(0)] * n
    my = [Fr(0)] * n
    sxx = [Fr(0)] * n
    syy = [Fr(0)] * n
    sxy = [Fr(0)] * n
    hess = {}
    frags = []
    for ti, t in enumerate(tri):
        p = [V[k] for k in t]
        if area(*p) < 0:
            p = [p[0], p[2], p[1]]
            fv = [F[t[0]], F[t[2]], F[t[1]]]
        else:
            fv = [F[k] for k in t]
        A2 = area(*p) * 2
        # linear density on the triangle via barycentric coordinates
        def rho(q, p=p, fv=fv, A2=A2):
            l1 = ((p[1][0] - q[0]) * (p[2][1] - q[1]) - (p[2][0] - q[0]) * (p[1][1] - q[1])) / A2
            l2 = ((p[2][0] - q[0]) * (p[0][1] - q[1]) - (p[0][0] - q[0]) * (p[2][1] - q[1])) / A2
            return l1 * fv[0] + l2 * fv[1] + (1 - l1 - l2) * fv[2]
        for i in range(n):
            poly, labs = list(p), [-1, -2, -3]
            for j in range(n):
                if j == i:
                    continue
                poly, labs = clip(poly, labs, Y, W, i, j)
                if not poly:
                    break
            if len(poly) < 3:
                continue
            # the fragment's exact area and centroid (fragments of zero area
            # are not listed)
            fa = Fr(0)
            fcx = Fr(0)
            fcy = Fr(0)
            for k in range(1, len(poly) - 1):
                a, b, c = poly[0], poly[k], poly[k + 1]
                ak = area(a, b, c)
                fa += ak
                fcx += ak * (a[0] + b[0] + c[0]) / 3
                fcy += ak * (a[1] + b[1] + c[1]) / 3
            if fa > 0:
                frags.append([i, ti, float(fa), float(fcx / fa), float(fcy / fa)])
            yi = Y[i]
            for k in range(1, len(poly) - 1):
                a, b, c = poly[0], poly[k], poly[k + 1]
                g = ((a[0] + b[0] + c[0]) / 3, (a[1] + b[1] + c[1]) / 3)
                mass[i] += area(a, b, c) * rho(g)
                intg[i] += ac6(a, b, c, lambda q: rho(q) * ((q[0] - yi[0]) ** 2 + (q[1] - yi[1]) ** 2))
                mx[i] += ac6(a, b, c, lambda q: rho(q) * q[0])
                my[i] += ac6(a, b, c, lambda q: rho(q) * q[1])
                sxx[i] += ac6(a, b, c, lambda q: rho(q) * q[0] * q[0])
                syy[i] += ac6(a, b, c, lambda q: rho(q) * q[1] * q[1])
                sxy[i] += ac6(a, b, c, lambda q: rho(q) * q[0] * q[1])
            for k in range(len(poly)):
                j = labs[k]
                if j < 0:
                    continue
                a, b = poly[k], poly[(k + 1) % len(poly)]
                l2 = (b[0] - a[0]) ** 2 + (b[1] - a[1]) ** 2
                mid = ((a[0] + b[0]) / 2, (a[1] + b[1]) / 2)
                d2 = (Y[i][0] - Y[j][0]) ** 2 + (Y[i][1] - Y[j][1]) ** 2
                val = mpmath.mpf(rho(mid).numerator) / rho(mid).denominator * \
                    mpmath.sqrt(mpmath.mpf(l2.numerator) / l2.denominator) / \
                    (2 * mpmath.sqrt(mpmath.mpf(d2.numerator) / d2.denominator))
                hess[(i, j)] = hess.get((i, j), 0) - val
                hess[(i, i)] = hess.get((i, i), 0) + val
    fval = sum(W[i] * mass[i] - intg[i] for i in range(n))
    return dict(
        fval=float(fval),
        mass=[float(m) for m in mass],
        intg=[float(v) for v in intg],
        mom=[[float(mx[i]), float(my[i])] for i in range(n)],
        inertia=[[float(sxx[i]), float(syy[i]), float(sxy[i])] for i in range(n)],
        total=float(sum(mass)),
        hessian=[[i, j, float(v)] for (i, j), v in sorted(hess.items())],
        fragments=sorted(frags),
    )


def case(name, vxy, vval, tri, X, w, note):
    out = solve(vxy, vval, tri, X, w)
    d = dict(name=name, note=note, vxy=vxy, vval=list(vval), tri=tri,
             X=[list(map(float, x)) for x in X], w=list(map(float, w)), expected=out)
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(d, f, indent=1)
    print(name, "mass sum", out["total"], "nnz", len(out["hessian"]))


def main():
    rng = np.random.Generator(np.random.PCG64(20240611))
    vxy, tri = grid(3)
    case("tiny_random", vxy, rng.uniform(0.5, 1.5, len(vxy)).tolist(), tri,
         rng.uniform(0, 1, (7, 2)).tolist(), rng.uniform(-0.01, 0.01, 7).tolist(),
         "3x3 grid, 7 random Diracs, small random weights")
    vxy, tri = grid(4)
    case("small_random", vxy, rng.uniform(0.5, 1.5, len(vxy)).tolist(), tri,
         rng.uniform(0, 1, (16, 2)).tolist(), [0.0] * 16,
         "4x4 grid, 16 random Diracs, zero weights (Voronoi)")
    vxy, tri = grid(4)
    Xw = rng.uniform(0, 1, (12, 2)).tolist()
    ww = rng.uniform(-0.05, 0.05, 12).tolist()
    ww[3] = -0.5  # a strongly negative weight: this Dirac's cell is empty
    case("weighted_hidden", vxy, rng.uniform(0.5, 1.5, len(vxy)).tolist(), tri, Xw, ww,
         "4x4 grid, 12 Diracs with weights in [-0.05,0.05]; Dirac 3 has weight -0.5 (empty cell)")
    vxy, tri = grid(3)
    Xo = rng.uniform(0, 1, (6, 2)).tolist()
    Xo[0] = [1.3, 0.4]
    Xo[1] = [-0.2, -0.3]
    case("outside", vxy, rng.uniform(0.5, 1.5, len(vxy)).tolist(), tri, Xo,
         rng.uniform(-0.01, 0.01, 6).tolist(), "3x3 grid, 6 Diracs, two outside the domain")
    # exact ties: a 4x4 lattice at cell centres on a 5x5 grid, uniform density,
    # zero weights: every bisector runs through mesh vertices
    vxy, tri = grid(5)
    Xl = [[(2 * i + 1) / 8.0, (2 * j + 1) / 8.0] for i in range(4) for j in range(4)]
    case("lattice_ties", vxy, [1.0] * len(vxy), tri, Xl, [0.0] * 16,
         "5x5 grid, uniform density 1, 4x4 lattice of Diracs at cell centres, zero weights")
    # non-convex L-shaped domain: 5x5 grid without the upper-right quarter
    vxy, tri = grid(5)
    keep = []
    for t in tri:
        cx = sum(vxy[k][0] for k in t) / 3
        cy = sum(vxy[k][1] for k in t) / 3
        if not (cx > 0.5 and cy > 0.5):
            keep.append(t)
    case("l_shape", vxy, rng.uniform(0.5, 1.5, len(vxy)).tolist(), keep,
         rng.uniform(0, 1, (10, 2)).tolist(), rng.uniform(-0.01, 0.01, 10).tolist(),
         "L-shaped domain (5x5 grid minus the upper-right quarter), 10 Diracs")
    vxy, tri = grid(3)
    case("single", vxy, rng.uniform(0.5, 1.5, len(vxy)).tolist(), tri, [[0.3, 0.6]], [0.0],
         "one Dirac: mass = total mass, empty Hessian")
    case("two", vxy, [1.0] * len(vxy), tri, [[0.25, 0.5], [0.75, 0.5]], [0.0, 0.02],
         "two Diracs, uniform density: bisector offset (w1-w0)/(2d)")


if __name__ == "__main__":
    main()
