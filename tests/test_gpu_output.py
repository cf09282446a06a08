"""GPU parity of the output entry points through the C ABI: fragment polygons
(ma_fragments, include/MA/voronoi_triangulation_intersection.hpp:315-343),
the rasterised Laguerre diagram (ma_draw_laguerre,
include/MA/rasterization.hpp:491-547) and the image density mesh
(ma_set_mesh_image, include/MA/functions.hpp:82-120).

Fragments: the exact-rational areas and centroids of the golden fixtures
(1e-12), and the FP64 oracle's polygons on seeded inputs (same (Dirac,
triangle) set, areas and centroids to 1e-12 of the cell scale), on every
cell path.  Images: the oracle's exact-coverage rasterisation to 1e-10 of
the largest pixel.  Image meshes: the same Kantorovich evaluation as the
host-built mesh of the same image, to the rounding of the sums.
"""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_CASES, frag_summary, load_golden
from mongeampere_amd import _lib, configs

pytestmark = pytest.mark.gpu

PATHS = {"fast": 0, "general": 1, "huge": 2}


def _compare_vertices(gpu, ref, scale):
    """Same polygon vertex by vertex: equal counts, and each vertex within
    1e-10 scale of one of the other's (the first vertex may differ)."""
    gk = {(int(gpu[0][f]), int(gpu[1][f])): gpu[3][gpu[2][f]:gpu[2][f + 1]]
          for f in range(len(gpu[0]))}
    rk = {(int(ref[0][f]), int(ref[1][f])): ref[3][ref[2][f]:ref[2][f + 1]]
          for f in range(len(ref[0]))}
    for k in set(gk) & set(rk):
        a, b = gk[k], rk[k]
        assert len(a) == len(b), (k, a, b)
        d = np.abs(a[:, None, :] - b[None, :, :]).max(axis=2)
        assert d.min(axis=1).max() <= 1e-10 * scale, (k, a, b)


def _compare_frags(got, want, scale, tiny=1e-12):
    """Same fragments (those of area below tiny * scale^2 may be missing on
    either side: exact ties), areas to 1e-12 scale^2, centroids to 1e-12
    scale."""
    for k in set(got) ^ set(want):
        a = got[k][0] if k in got else want[k][0]
        assert abs(a) <= tiny * scale * scale, f"fragment {k} of area {a} on one side only"
    for k in set(got) & set(want):
        ga, gx, gy = got[k]
        wa, wx, wy = want[k]
        assert abs(ga - wa) <= 1e-12 * scale * scale, (k, ga, wa)
        if abs(wa) > 1e-9 * scale * scale:
            assert abs(gx - wx) <= 1e-10 * scale and abs(gy - wy) <= 1e-10 * scale, (k, gx, wx)


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fragments_golden(engine, name, path):
    c = load_golden(name)
    print(f"case {name} path {path}", flush=True)
    engine.set_option("force_general", PATHS[path])
    try:
        engine.set_mesh(c["mesh"]).set_diracs(c["X"])
        cell, tri, start, xy = engine.fragments(c["w"])
    finally:
        engine.set_option("force_general", 0)
    # sorted by (Dirac, triangle), counter-clockwise
    key = cell.astype(np.int64) * (1 << 32) + tri
    assert np.all(np.diff(key) > 0)
    got = frag_summary(cell, tri, start, xy)
    assert all(v[0] > 0 for v in got.values())
    want = {(int(r[0]), int(r[1])): tuple(r[2:]) for r in c["frags"]}
    _compare_frags(got, want, 1.0)


def _random_case(seed, N, n):
    rng = np.random.Generator(np.random.PCG64(2000 + seed))
    vxy, tri = configs.grid_mesh(n)
    vval = rng.uniform(0.5, 1.5, vxy.shape[0])
    X = rng.uniform(0, 1, (N, 2))
    w = rng.uniform(-0.3, 0.3, N) / N
    return configs.Mesh(vxy, vval, tri), X, w


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("seed,N,n", [(0, 50, 6), (1, 2000, 40), (2, 20000, 64)])
def test_fragments_vs_oracle(engine, seed, N, n, path):
    mesh, X, w = _random_case(seed, N, n)
    if path != "fast" and N > 2000:
        pytest.skip("the one-thread and one-workgroup paths are covered at N <= 2000")
    print(f"seed {seed} N {N} path {path}", flush=True)
    engine.set_option("force_general", PATHS[path])
    try:
        engine.set_mesh(mesh).set_diracs(X)
        cell, tri, start, xy = engine.fragments(w)
    finally:
        engine.set_option("force_general", 0)
    ocell, otri, ostart, oxy = oracle.fragments(mesh, X, w)
    _compare_frags(frag_summary(cell, tri, start, xy), frag_summary(ocell, otri, ostart, oxy),
                   1.0 / np.sqrt(N))
    _compare_vertices((cell, tri, start, xy), (ocell, otri, ostart, oxy), 1.0 / np.sqrt(N))
    # fragment masses add up to the first-moment masses
    m, _ = engine.first_moment(w)
    T, V, F = mesh.tri, mesh.vxy, mesh.vval
    got = np.zeros(N)
    for (i, t), (a, cx, cy) in frag_summary(cell, tri, start, xy).items():
        P = V[T[t]]
        A = np.array([[P[0, 0], P[0, 1], 1], [P[1, 0], P[1, 1], 1], [P[2, 0], P[2, 1], 1]])
        coef = np.linalg.solve(A, F[T[t]])
        got[i] += a * (coef[0] * cx + coef[1] * cy + coef[2])
    assert np.abs(got - m).max() <= 1e-10 * m.max()


def test_fragments_capacity_protocol(engine):
    mesh, X, w = _random_case(3, 300, 12)
    engine.set_mesh(mesh).set_diracs(X)
    L = _lib.lib()
    nf = ctypes.c_int64()
    nv = ctypes.c_int64()
    wp = np.ascontiguousarray(w)
    st = L.ma_fragments(engine._h, wp.ctypes.data, 0, 0, ctypes.byref(nf), ctypes.byref(nv),
                        None, None, None, None)
    assert st == _lib.MA_CAPACITY and nf.value > 300 and nv.value >= 3 * nf.value
    cell = np.empty(nf.value, np.int32)
    tri = np.empty(nf.value, np.int32)
    start = np.empty(nf.value + 1, np.int64)
    xy = np.empty((nv.value, 2))
    # one vertex short: refused, nothing written
    st = L.ma_fragments(engine._h, wp.ctypes.data, nf.value, nv.value - 1, ctypes.byref(nf),
                        ctypes.byref(nv), cell.ctypes.data, tri.ctypes.data,
                        start.ctypes.data, xy.ctypes.data)
    assert st == _lib.MA_CAPACITY
    st = L.ma_fragments(engine._h, wp.ctypes.data, nf.value, nv.value, ctypes.byref(nf),
                        ctypes.byref(nv), cell.ctypes.data, tri.ctypes.data,
                        start.ctypes.data, xy.ctypes.data)
    assert st == 0 and start[0] == 0 and start[-1] == nv.value
    # the fragments tile the domain
    areas = np.array([v[0] for v in frag_summary(cell, tri, start, xy).values()])
    assert np.isclose(areas.sum(), 1.0, rtol=1e-13)


@pytest.mark.parametrize("path", list(PATHS))
def test_draw_laguerre_vs_oracle(engine, path):
    mesh, X, w = _random_case(4, 400, 16)
    rng = np.random.default_rng(5)
    colors = rng.uniform(0, 1, (400, 3))
    box, W, H = (-0.05, 0.02, 1.1, 0.97), 37, 29
    print(f"draw path {path}", flush=True)
    engine.set_option("force_general", PATHS[path])
    try:
        engine.set_mesh(mesh).set_diracs(X)
        img = engine.draw_laguerre_diagram(w, colors, box, W, H)
    finally:
        engine.set_option("force_general", 0)
    ref = oracle.draw_laguerre_diagram(mesh, X, w, colors, box, W, H)
    assert img.shape == ref.shape
    assert np.abs(img - ref).max() <= 1e-10 * np.abs(ref).max()


def _hip():
    """The HIP runtime the engine links, for device buffers in tests."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _DevArray:
    """A device copy of a float64 array (hipMalloc + hipMemcpy)."""

    def __init__(self, hip, a):
        self.hip = hip
        self.a = np.ascontiguousarray(a, np.float64)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), self.a.nbytes) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_draw_laguerre_large_fragments_and_device_pointers(engine):
    # few cells on a fine image: fragments cover many pixels (workgroup kernel)
    mesh, X, w = _random_case(6, 20, 4)
    colors = np.random.default_rng(7).uniform(0, 1, (20, 2))
    box, W, H = (0.0, 0.0, 1.0, 1.0), 160, 120
    engine.set_mesh(mesh).set_diracs(X)
    img = engine.draw_laguerre_diagram(w, colors, box, W, H)
    ref = oracle.draw_laguerre_diagram(mesh, X, w, colors, box, W, H)
    assert np.abs(img - ref).max() <= 1e-10 * np.abs(ref).max()
    hip = _hip()
    dw, dc, di = _DevArray(hip, w), _DevArray(hip, colors), _DevArray(hip, np.zeros((H, W, 2)))
    try:
        engine.draw_laguerre_device(dw.p.value, dc.p.value, 2, box, W, H, di.p.value)
        assert np.abs(di.get() - ref).max() <= 1e-10 * np.abs(ref).max()
    finally:
        for d in (dw, dc, di):
            d.free()


@pytest.mark.parametrize("shape", [(5, 7), (33, 20), (128, 96)])
def test_set_mesh_image_matches_host_mesh(engine, shape):
    from mongeampere_amd.api import Engine

    H, W = shape
    rng = np.random.default_rng(H * W)
    img = rng.integers(0, 256, (H, W)).astype(np.float64)
    mesh, total_host = configs.image_to_pl(img)
    N = max(H * W // 8, 4)
    X = rng.uniform(-1, 1, (N, 2))
    w = rng.uniform(-1e-3, 1e-3, N)
    print(f"image {shape}", flush=True)
    total = engine.set_mesh_image(img)
    assert abs(total - total_host) <= 1e-13 * total_host
    engine.set_diracs(X)
    f1, g1, H1 = engine.kantorovich(w)
    with Engine(0) as e2:
        e2.set_mesh(mesh).set_diracs(X)
        f2, g2, H2 = e2.kantorovich(w)
    # same mesh: equal up to the order of the floating-point sums
    assert abs(f1 - f2) <= 1e-14 * abs(f2)
    assert np.abs(g1 - g2).max() <= 1e-14 * np.abs(g2).max()
    assert abs(H1 - H2).max() <= 1e-14 * abs(H2).max()


def _poly_area(q):
    return 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])


def _convex_polygon(seed):
    from scipy.spatial import ConvexHull

    pts = np.random.default_rng(seed).uniform(-1, 1, (12, 2))
    return pts[ConvexHull(pts).vertices]


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("shape", ["square", "hexagon", "random"])
def test_cell_polygons_vs_oracle(engine, shape, path):
    if shape == "square":
        P = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
    elif shape == "hexagon":
        t = np.linspace(0, 2 * np.pi, 7)[:-1]
        P = np.stack([np.cos(t), np.sin(t)], 1)[::-1]  # clockwise input
    else:
        P = _convex_polygon(11)
    rng = np.random.default_rng(len(shape))
    N = 60
    X = rng.uniform(-0.8, 0.8, (N, 2))
    w = rng.uniform(-0.02, 0.02, N)
    print(f"polygon {shape} path {path}", flush=True)
    engine.set_option("force_general", PATHS[path])
    try:
        engine.set_mesh_polygon(P).set_diracs(X)
        start, xy = engine.cell_polygons(w)
    finally:
        engine.set_option("force_general", 0)
    ref = oracle.voronoi_polygon_intersection(P, X, w)
    for i in range(N):
        q = xy[start[i]:start[i + 1]]
        r = ref[i]
        assert len(q) == len(r), (i, q, r)
        if len(r):
            d = np.abs(q[:, None, :] - r[None, :, :]).max(axis=2)
            assert d.min(axis=1).max() <= 1e-12, (i, q, r)
            assert _poly_area(q) > 0
    total = sum(_poly_area(xy[start[i]:start[i + 1]]) for i in range(N) if start[i + 1] > start[i])
    assert np.isclose(total, abs(_poly_area(P)), rtol=1e-13)


def test_cell_polygons_lattice_and_large(engine):
    from mongeampere_amd.api import voronoi_polygon_intersection

    SQ = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
    k = 7
    c = -1 + (2 * np.arange(k) + 1) / k
    X = np.stack(np.meshgrid(c, c, indexing="ij"), -1).reshape(-1, 2)
    cells = voronoi_polygon_intersection(SQ, X, np.zeros(k * k), engine=engine)
    for i, q in enumerate(cells):
        assert q.shape == (4, 2)
        assert abs(_poly_area(q) - (2.0 / k) ** 2) <= 1e-14
        assert np.abs(q.mean(axis=0) - X[i]).max() <= 1e-14
    # the reference's test_voronoi / test_power workload at N = 100000:
    # random Diracs on [-1,1]^2, the cells tile the square (area 4)
    X = np.random.default_rng(3).uniform(-1, 1, (100000, 2))
    engine.set_mesh_polygon(SQ).set_diracs(X)
    start, xy = engine.cell_polygons(np.zeros(100000))
    cnt = np.diff(start)
    assert cnt.min() >= 3
    idx = np.repeat(np.arange(100000), cnt)
    nxt = np.arange(xy.shape[0]) + 1
    nxt[start[1:] - 1] = start[:-1]
    cross = xy[:, 0] * xy[nxt, 1] - xy[nxt, 0] * xy[:, 1]
    areas = 0.5 * np.bincount(idx, cross, minlength=100000)
    assert areas.min() > 0
    assert abs(areas.sum() - 4.0) <= 1e-11
    m, _ = engine.first_moment(np.zeros(100000))
    # (shoelace areas in absolute coordinates carry ~1e-16 of rounding)
    assert np.abs(areas - m).max() <= 1e-11 * m.max()


def test_output_entry_points_reject_bad_input(engine):
    from mongeampere_amd._lib import EngineError

    # a non-convex polygon, a repeated vertex and a degenerate box are refused
    arrow = np.array([[0.0, 0.0], [1.0, 0.0], [0.3, 0.3], [0.0, 1.0]])
    with pytest.raises(EngineError):
        engine.set_mesh_polygon(arrow)
    with pytest.raises(EngineError):
        engine.set_mesh_polygon(np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0], [0.0, 1.0]]))
    mesh, X, w = _random_case(8, 30, 6)
    engine.set_mesh(mesh).set_diracs(X)
    with pytest.raises(EngineError):
        engine.draw_laguerre_diagram(w, np.ones(30), (1.0, 0.0, 0.0, 1.0), 8, 8)
    with pytest.raises(ValueError):
        engine.draw_laguerre_diagram(w, np.ones(29), (0.0, 0.0, 1.0, 1.0), 8, 8)
    with pytest.raises(EngineError):
        engine.set_mesh_image(np.ones((1, 5)))
    # a Dirac hidden by a heavier coincident one: no fragments, no polygon
    X2 = np.vstack([X, X[:1]])
    w2 = np.append(w, w[0] + 1e-3)
    engine.set_diracs(X2)
    cell, tri, start, xy = engine.fragments(w2)
    assert 0 not in set(cell.tolist()) and 30 in set(cell.tolist())
    P = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    engine.set_mesh_polygon(P).set_diracs(X2)
    s, q = engine.cell_polygons(w2)
    assert s[1] == s[0] and s[31] > s[30]
