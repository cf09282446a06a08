"""Mixed precision (FP32 cell geometry and quadrature, FP64 decisions near
zero, FP64 damping / CG / final check) through the C ABI.

Stated tolerances of an FP32 evaluation against the FP64 oracle:
  masses     max_i |m_i - m_i^oracle| <= 2e-5 * mean(m)
  value      |f - f^oracle| <= 1e-5 * |f^oracle|
  Hessian    max |H - H^oracle| <= 1e-4 * max |H^oracle|
A mixed solve must reach the same FP64 stopping test as the FP64 solve.
"""
import numpy as np
import pytest

import oracle
from mongeampere_amd import configs

pytestmark = pytest.mark.gpu

MASS_TOL, F_TOL, H_TOL = 2e-5, 1e-5, 1e-4


def _check_eval(engine, p, w):
    engine.set_mesh(p.mesh).set_diracs(p.X)
    f32, g32, H32 = engine.kantorovich(w, precision="mixed")
    info = engine.eval_info()
    of, og, oH = oracle.kantorovich(p.mesh, p.X, w)
    scale = np.mean(og)
    assert np.max(np.abs(g32 - og)) <= MASS_TOL * scale
    assert abs(f32 - of) <= F_TOL * abs(of)
    D = (H32 - oH).tocoo()
    hmax = np.max(np.abs(oH.data))
    assert (np.max(np.abs(D.data)) if D.nnz else 0.0) <= H_TOL * hmax
    assert abs(g32.sum() - og.sum()) <= 1e-6 * og.sum()
    return info


@pytest.mark.parametrize("wscale", [0.0, 1.0])
def test_mixed_eval_config1(engine, wscale):
    p = configs.config1()
    rng = np.random.default_rng(21)
    w = wscale * rng.uniform(-0.3, 0.3, p.X.shape[0]) / p.X.shape[0]
    info = _check_eval(engine, p, w)
    assert info["nredecide"] >= 0


@pytest.mark.parametrize("seed", range(4))
def test_mixed_eval_random(engine, seed):
    p = configs.config1(N=300 + 100 * seed, n=17 + 4 * seed)
    rng = np.random.default_rng(100 + seed)
    w = rng.uniform(-0.5, 0.5, p.X.shape[0]) / p.X.shape[0]
    _check_eval(engine, p, w)


def test_mixed_eval_lattice_ties(engine):
    """Dyadic lattice at w = 0: every cell edge lies on a tie; the FP32 path
    must still give the closed-form masses 1/N."""
    k = 16
    h = 1.0 / k
    g = (np.arange(k) + 0.5) * h
    X = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    vxy, tri = configs.grid_mesh(9)
    mesh = configs.Mesh(vxy, np.ones(vxy.shape[0]), tri)
    engine.set_mesh(mesh).set_diracs(X)
    f32, g32, H32 = engine.kantorovich(np.zeros(len(X)), precision="mixed")
    assert np.max(np.abs(g32 - 1.0 / len(X))) <= MASS_TOL / len(X)


def test_mixed_solve_config1(engine):
    p = configs.config1()
    engine.set_mesh(p.mesh).set_diracs(p.X)
    w32, st32 = engine.ot_solve(p.nu, stop="maxrel", tol=1e-9, precision="mixed")
    assert st32.status == "MA_OK"
    assert st32.max_rel_err <= 1e-9
    assert st32.neval_fp32 > 0
    w64, st64 = engine.ot_solve(p.nu, stop="maxrel", tol=1e-9, precision="fp64")
    assert np.max(np.abs(w32 - w64)) <= 1e-6 * np.max(np.abs(w64))
    _, og, _ = oracle.kantorovich(p.mesh, p.X, w32)
    assert np.max(np.abs(og - p.nu) / p.nu) <= 2e-9


def test_mixed_solve_reference_rule(engine):
    """With the reference's absolute L2 rule the solve may only stop on an
    FP64 evaluation."""
    p = configs.config1(N=400, n=24)
    engine.set_mesh(p.mesh).set_diracs(p.X)
    w, st = engine.ot_solve(p.nu, eps_g=1e-7, precision="mixed")
    assert st.status == "MA_OK"
    _, og, _ = oracle.kantorovich(p.mesh, p.X, w)
    assert np.linalg.norm(og - p.nu) < 1.5e-7


def test_mixed_config3_scaled(engine):
    """C3-like density at reduced size: FP32 vs FP64 GPU evaluations."""
    p = configs.config3(N=50_000, n=256)
    engine.set_mesh(p.mesh).set_diracs(p.X)
    rng = np.random.default_rng(5)
    w = rng.uniform(-0.2, 0.2, p.X.shape[0]) / p.X.shape[0]
    f64, g64, H64 = engine.kantorovich(w, precision="fp64")
    f32, g32, H32 = engine.kantorovich(w, precision="mixed")
    assert np.max(np.abs(g32 - g64)) <= MASS_TOL * np.mean(g64)
    assert abs(f32 - f64) <= F_TOL * abs(f64)
    D = (H32 - H64).tocoo()
    assert (np.max(np.abs(D.data)) if D.nnz else 0.0) <= H_TOL * np.max(np.abs(H64.data))
