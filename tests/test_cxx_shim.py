"""The header-only C++ shim (include/MA/*.hpp) keeps the reference's call
shapes (kantorovich, ot_solve, solve_laplacian_matrix, lloyd, second_moment,
voronoi_triangulation_intersection, draw_laguerre_diagram).  CPU: it compiles and links against stand-ins of the CGAL / Eigen
types.  GPU: the prebuilt driver program gives the same results as the
Python binding of the C ABI on the same inputs."""
import json
import os
import subprocess

import numpy as np
import pytest

from mongeampere_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "shim_mock.cpp")
BIN = os.path.join(ROOT, "tests", "cxx", "shim_mock")
LIBDIR = os.path.join(ROOT, "mongeampere_amd")


def test_shim_compiles_against_mock_types(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libmagpu.so")):
        pytest.skip("engine library not built")
    out = tmp_path / "shim_mock"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR, "-lmagpu",
                        "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_input(path, p):
    with open(path, "w") as f:
        f.write(f"{p.mesh.vxy.shape[0]}\n")
        for (x, y), v in zip(p.mesh.vxy.tolist(), p.mesh.vval.tolist()):
            f.write(f"{x!r} {y!r} {v!r}\n")
        f.write(f"{p.mesh.tri.shape[0]}\n")
        for a, b, c in p.mesh.tri.tolist():
            f.write(f"{a} {b} {c}\n")
        f.write(f"{p.X.shape[0]}\n")
        for (x, y), nu in zip(p.X.tolist(), p.nu.tolist()):
            f.write(f"{x!r} {y!r} {nu!r}\n")


@pytest.mark.gpu
def test_shim_matches_c_abi(engine, tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/shim_mock was not built (run __graft_entry__.build())")
    p = configs.config1(N=300, n=17)
    inp = tmp_path / "in.txt"
    _write_input(inp, p)
    r = subprocess.run([BIN, str(inp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    engine.set_mesh(p.mesh).set_diracs(p.X)
    N = p.X.shape[0]
    f, g, H = engine.kantorovich(np.zeros(N))
    assert abs(out["fval"] - f) <= 1e-12 * abs(f)
    assert np.max(np.abs(np.array(out["g"]) - g)) <= 1e-14
    Hs = np.zeros((N, N))
    for i, j, v in out["hessian"]:
        Hs[i, j] += v
    assert np.max(np.abs(Hs - H.toarray())) <= 1e-12 * np.max(np.abs(H.data))
    d, _, _ = engine.solve_laplacian(g - p.nu, rel_tol=1e-12)
    assert np.max(np.abs(np.array(out["d"]) - d)) <= 1e-8 * np.max(np.abs(d))
    w, st = engine.ot_solve(p.nu, eps_g=1e-10)
    assert np.max(np.abs(np.array(out["w"]) - w)) <= 1e-8 * np.max(np.abs(w))
    assert out["niter"] == st.niter and out["neval"] == st.neval
    m, c = engine.lloyd(w)
    assert np.max(np.abs(np.array(out["centroids"]) - c)) <= 1e-12
    _, _, inertia = engine.second_moment(w)
    assert np.max(np.abs(np.array(out["inertia"]) - inertia)) <= 1e-14
    # traversal callback and rasterisation at the driver's solved weights
    wm = np.array(out["w"])
    cell, tri, start, xy = engine.fragments(wm)
    assert out["nfrag"] == cell.shape[0]
    area = np.zeros(N)
    for f in range(cell.shape[0]):
        q = xy[start[f]:start[f + 1]]
        area[cell[f]] += 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])
    assert np.max(np.abs(np.array(out["frag_area"]) - area)) <= 1e-14
    img = engine.draw_laguerre_diagram(wm, np.arange(1.0, N + 1.0), (0, 0, 1, 1), 16, 12)
    assert np.max(np.abs(np.array(out["image"]) - img[..., 0].ravel())) <= 1e-12 * img.max()
