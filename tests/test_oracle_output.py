"""CPU checks of the oracle's output functions: fragment polygons
(include/MA/voronoi_triangulation_intersection.hpp:315-343), the rasterised
Laguerre diagram (include/MA/rasterization.hpp:403-547) and the P1 density of
an image (include/MA/functions.hpp:82-120).

Pins: the exact-rational fragment areas and centroids of the golden fixtures;
closed forms of the rasterisation (a single cell covering the box, a vertical
bisector splitting pixels); brute-force coverage by point sampling; the
image mesh's total mass for constant and linear images (the P1 interpolant of
a linear function is exact)."""
import numpy as np
import pytest

import oracle
from conftest import GOLDEN_CASES, frag_summary, load_golden
from mongeampere_amd import configs


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fragments_vs_golden(name):
    g = load_golden(name)
    cell, tri, start, xy = oracle.fragments(g["mesh"], g["X"], g["w"])
    got = frag_summary(cell, tri, start, xy)
    want = {(int(r[0]), int(r[1])): r[2:] for r in g["frags"]}
    assert set(got) == set(want)
    for k, (a, cx, cy) in want.items():
        ga, gx, gy = got[k]
        assert abs(ga - a) <= 1e-13 * max(1.0, abs(a)), (k, ga, a)
        assert abs(gx - cx) <= 1e-12 and abs(gy - cy) <= 1e-12, (k, gx, cx, gy, cy)


def test_fragments_sum_to_first_moment_masses():
    p = configs.config1(N=200, n=12)
    w = np.random.default_rng(3).uniform(-1e-3, 1e-3, 200)
    cell, tri, start, xy = oracle.fragments(p.mesh, p.X, w)
    m, _ = oracle.first_moment(p.mesh, p.X, w)
    # mass of a fragment = area x density at its centroid (centroid rule,
    # exact for the linear density of its triangle)
    T = p.mesh.tri
    V = p.mesh.vxy
    F = p.mesh.vval
    got = np.zeros(200)
    for (i, t), (a, cx, cy) in frag_summary(cell, tri, start, xy).items():
        P = V[T[t]]
        A = np.array([[P[0, 0], P[0, 1], 1], [P[1, 0], P[1, 1], 1], [P[2, 0], P[2, 1], 1]])
        coef = np.linalg.solve(A, F[T[t]])
        got[i] += a * (coef[0] * cx + coef[1] * cy + coef[2])
    assert np.allclose(got, m, rtol=1e-11, atol=1e-15)


def _unit_square_mesh(vals):
    vxy, tri = configs.grid_mesh(2)
    return configs.Mesh(vxy, np.asarray(vals, np.float64), tri)


def test_draw_single_cell_covers_everything():
    mesh = _unit_square_mesh(np.ones(4))
    img = oracle.draw_laguerre_diagram(mesh, np.array([[0.3, 0.6]]), np.zeros(1),
                                       np.array([[2.0, -1.0]]), (0, 0, 1, 1), 7, 5)
    # each pixel is covered by the two triangles' fragments, coverage 1 in all
    assert np.allclose(img[..., 0], 2.0, atol=1e-13)
    assert np.allclose(img[..., 1], -1.0, atol=1e-13)


@pytest.mark.parametrize("W", [3, 4, 5])
def test_draw_vertical_bisector(W):
    mesh = _unit_square_mesh(np.ones(4))
    X = np.array([[0.25, 0.5], [0.75, 0.5]])
    img = oracle.draw_laguerre_diagram(mesh, X, np.zeros(2), np.array([1.0, 10.0]),
                                       (0, 0, 1, 1), W, 2)[..., 0]
    # bisector x = 0.5 at pixel coordinate W/2
    want = np.empty(W)
    for i in range(W):
        left = min(max(W / 2 - i, 0.0), 1.0)
        want[i] = left * 1.0 + (1 - left) * 10.0
    assert np.allclose(img, want[None, :], atol=1e-13)


def test_draw_coverage_vs_point_sampling():
    p = configs.config1(N=12, n=4)
    w = np.zeros(12)
    box = (-0.1, -0.05, 1.05, 1.1)
    W, H = 9, 7
    img = oracle.draw_laguerre_diagram(p.mesh, p.X, w, np.ones(12), box, W, H)[..., 0]
    # same picture by sampling: each fragment's colour x mean vertex density
    # on a 64 x 64 grid of points per pixel
    cell, tri, start, xy = oracle.fragments(p.mesh, p.X, w)
    T, V, F = p.mesh.tri, p.mesh.vxy, p.mesh.vval
    s = 64
    ref = np.zeros((H, W))
    for f in range(len(cell)):
        poly = xy[start[f]:start[f + 1]]
        P = V[T[tri[f]]]
        A = np.array([[P[0, 0], P[0, 1], 1], [P[1, 0], P[1, 1], 1], [P[2, 0], P[2, 1], 1]])
        coef = np.linalg.solve(A, F[T[tri[f]]])
        ff = np.mean(poly @ coef[:2] + coef[2])
        u = W * (poly[:, 0] - box[0]) / (box[2] - box[0])
        v = H * (poly[:, 1] - box[1]) / (box[3] - box[1])
        for j in range(max(int(v.min()), 0), min(int(v.max()) + 1, H)):
            for i in range(max(int(u.min()), 0), min(int(u.max()) + 1, W)):
                t = (np.arange(s) + 0.5) / s
                qu, qv = np.meshgrid(i + t, j + t)
                inside = np.ones(qu.shape, bool)
                for k in range(len(u)):
                    k1 = (k + 1) % len(u)
                    cr = (u[k1] - u[k]) * (qv - v[k]) - (v[k1] - v[k]) * (qu - u[k])
                    inside &= cr >= 0
                ref[j, i] += inside.mean() * ff
    assert np.abs(img - ref).max() < 0.03 * img.max()
    # and the total: sum of fragment areas (pixel units) x weights
    assert np.isclose(img.sum(), ref.sum(), rtol=5e-3)


def test_image_mesh_constant_and_linear():
    W, H = 7, 5
    img = np.full((H, W), 51.0)
    mesh, total = configs.image_to_pl(img)
    assert np.isclose(total, 4 * (51.0 / 255 + 1e-3), rtol=1e-14)
    # pixel value = column index: the density is linear in x, its P1
    # interpolant exact, the integral 4 x the mean
    ramp = np.tile(np.arange(W, dtype=np.float64), (H, 1))
    mesh, total = configs.image_to_pl(ramp)
    assert np.isclose(total, 4 * ((W - 1) / 2 / 255 + 1e-3), rtol=1e-14)
    # the first image row is the top: pixel row 0 -> vertices with y = +1
    top = np.zeros((H, W))
    top[0, :] = 255.0
    mesh, _ = configs.image_to_pl(top)
    hi = mesh.vval > 0.5
    assert np.all(mesh.vxy[hi, 1] == 1.0) and hi.sum() == W
    assert mesh.vxy[1, 0] == -1.0 and mesh.vxy[1, 1] == -1.0 + 2.0 / (H - 1)


def _poly_area(q):
    return 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])


SQUARE = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_polygon_cells_lattice_closed_form(k):
    # centred k x k lattice on [-1,1]^2, w = 0: square cells of side 2/k
    c = -1 + (2 * np.arange(k) + 1) / k
    X = np.stack(np.meshgrid(c, c, indexing="ij"), -1).reshape(-1, 2)
    cells = oracle.voronoi_polygon_intersection(SQUARE, X, np.zeros(k * k))
    for i, q in enumerate(cells):
        assert q.shape == (4, 2)
        assert np.isclose(_poly_area(q), (2.0 / k) ** 2, rtol=1e-15)
        assert np.allclose(q.mean(axis=0), X[i], atol=1e-15)


def test_polygon_cells_two_diracs_and_total_area():
    # two Diracs on the x axis, weights (w0, w1): bisector at
    # x = (|y1|^2 - |y0|^2 - w1 + w0) / (2 (x1 - x0))
    X = np.array([[-0.5, 0.0], [0.5, 0.0]])
    w = np.array([0.1, -0.1])
    cells = oracle.voronoi_polygon_intersection(SQUARE, X, w)
    xb = (0.25 - 0.25 + 0.1 + 0.1) / 2.0
    assert np.isclose(_poly_area(cells[0]), 2 * (xb + 1), rtol=1e-15)
    assert np.isclose(_poly_area(cells[1]), 2 * (1 - xb), rtol=1e-15)
    # random Diracs on a hexagon: the cells tile it (tests/test_voronoi.cpp:50)
    t = np.linspace(0, 2 * np.pi, 7)[:-1]
    hexa = np.stack([np.cos(t), np.sin(t)], 1)
    rng = np.random.default_rng(1)
    X = rng.uniform(-0.7, 0.7, (40, 2))
    cells = oracle.voronoi_polygon_intersection(hexa, X, rng.uniform(-0.01, 0.01, 40))
    assert np.isclose(sum(_poly_area(q) for q in cells if len(q)), _poly_area(hexa), rtol=1e-13)
