"""GPU Newton solver parity (FP64 mode) through the C ABI."""
import numpy as np
import pytest

import oracle
from mongeampere_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precond", ["amg", "jacobi"])
def test_solve_config1_maxrel(engine, precond):
    """Config 1: damped Newton from w=0 to max |m-nu|/nu <= 1e-9; masses at
    the GPU solution checked with the oracle; weights equal the oracle's
    Newton solution modulo the gauge (the last weight is never moved)."""
    p = configs.config1()
    engine.set_mesh(p.mesh).set_diracs(p.X)
    w, st = engine.ot_solve(p.nu, stop="maxrel", tol=1e-9, preconditioner=precond)
    assert st.status == "MA_OK"
    assert st.max_rel_err <= 1e-9
    _, og, _ = oracle.kantorovich(p.mesh, p.X, w)
    assert np.max(np.abs(og - p.nu) / p.nu) <= 2e-9
    ow, ost = oracle.ot_solve(p.mesh, p.X, p.nu, stop="maxrel", tol=1e-9)
    assert ost.status == "ok"
    assert abs(w[-1]) == 0.0  # gauge: w_{N-1} stays at its initial value
    assert np.max(np.abs(w - ow)) <= 1e-6 * np.max(np.abs(ow))


def test_solve_reference_rule(engine):
    """The reference stopping rule ||m - nu||_2 < eps_g."""
    p = configs.config1(N=400, n=24)
    engine.set_mesh(p.mesh).set_diracs(p.X)
    w, st = engine.ot_solve(p.nu, eps_g=1e-7)
    assert st.status == "MA_OK"
    assert st.l2_err < 1e-7
    _, og, _ = oracle.kantorovich(p.mesh, p.X, w)
    assert np.linalg.norm(og - p.nu) < 1.5e-7


def test_solve_laplacian_vs_direct(engine):
    """Gauge-fixed linear solve vs the oracle's sparse direct solve
    (include/MA/optimal_transport.hpp:41-87)."""
    p = configs.config1()
    engine.set_mesh(p.mesh).set_diracs(p.X)
    _, g, H = engine.kantorovich(np.zeros(p.X.shape[0]))
    rhs = g - p.nu
    d, res, it = engine.solve_laplacian(rhs, rel_tol=1e-13)
    od, err = oracle.solve_laplacian_matrix(H, rhs)
    assert d[-1] == 0.0
    assert np.max(np.abs(d - od)) <= 1e-8 * np.max(np.abs(od))


def test_solve_laplacian_caller_csr(engine):
    """Caller-supplied CSR (solve_laplacian_matrix(h, g) call shape) in the
    caller's Dirac order: same solution as the direct solve."""
    p = configs.config1()
    engine.set_mesh(p.mesh).set_diracs(p.X)
    _, g, H = engine.kantorovich(np.zeros(p.X.shape[0]))
    rhs = g - p.nu
    d, res, it = engine.solve_laplacian_matrix(H, rhs, rel_tol=1e-13)
    od, err = oracle.solve_laplacian_matrix(H, rhs)
    assert d[-1] == 0.0
    assert np.max(np.abs(d - od)) <= 1e-8 * np.max(np.abs(od))


def test_graph_replay_matches_eager(engine):
    """CG batches replayed from a HIP graph give bit-identical results to
    eager launches, over several solves with eager work in between."""
    p = configs.config1()
    engine.set_mesh(p.mesh).set_diracs(p.X)
    try:
        for trial in range(3):
            # weights of the first Newton steps (every cell non-empty)
            w, _ = engine.ot_solve(p.nu, stop="maxrel", tol=1e-12, maxiter=trial - 1,
                                   raise_on_error=False) if trial else (np.zeros(p.X.shape[0]), None)
            _, g, _ = engine.kantorovich(w)  # one Hessian for both solves
            out = {}
            for graphs in (1, 0, 1):
                engine.set_option("graphs", graphs)
                d, _, it = engine.solve_laplacian(g - p.nu, rel_tol=1e-12)
                if graphs in out:
                    assert np.array_equal(d, out[graphs][0])  # second replay
                out[graphs] = (d, it)
            assert out[1][1] == out[0][1]
            assert np.array_equal(out[1][0], out[0][0])
    finally:
        engine.set_option("graphs", 1)


def test_empty_cell_status(engine):
    """A Dirac whose cell is empty at w=0 makes the solve stop with
    MA_EMPTY_CELL and report its index (reference: eps0 <= 0 early return,
    include/MA/optimal_transport.hpp:137-148)."""
    p = configs.config1(N=50, n=9)
    X = p.X.copy()
    X[7] = X[3]  # coincident Diracs: one of them is hidden
    engine.set_mesh(p.mesh).set_diracs(X)
    w, st = engine.ot_solve(p.nu, raise_on_error=False)
    assert st.status == "MA_EMPTY_CELL"
    assert st.bad_index in (3, 7)


def test_engine_reuse_across_sizes():
    """One engine solving problems of growing and shrinking size (both
    precisions) gives the same weights as a fresh engine each time: buffers
    exchanged by the Newton driver keep their capacities."""
    from mongeampere_amd.api import Engine
    probs = [configs.config1(), configs.config3(N=20_000, n=128), configs.config1(N=400, n=24),
             configs.config3(N=30_000, n=160)]
    shared = Engine(0)
    for i, p in enumerate(probs):
        for prec in ("mixed", "fp64"):
            print(f"problem {i} N={p.X.shape[0]} {prec}", flush=True)
            shared.set_mesh(p.mesh).set_diracs(p.X)
            w1, s1 = shared.ot_solve(p.nu, stop="maxrel", tol=1e-9, precision=prec)
            with Engine(0) as fresh:
                fresh.set_mesh(p.mesh).set_diracs(p.X)
                w2, s2 = fresh.ot_solve(p.nu, stop="maxrel", tol=1e-9, precision=prec)
            assert s1.status == s2.status == "MA_OK"
            assert np.max(np.abs(w1 - w2)) <= 1e-9 * np.max(np.abs(w2)) + 1e-15
    shared.close()


def test_multigrid_cycle_options(engine):
    """V(1,1) and V(2,2) cycles with other damping converge to the same
    weights (the preconditioner changes the path, not the solution); values
    outside the documented ranges are refused."""
    from mongeampere_amd._lib import EngineError

    p = configs.config2(N=20000, n=128)
    engine.set_mesh(p.mesh).set_diracs(p.X)
    ws = []
    try:
        for sweeps, omega in ((2, 850), (1, 850), (2, 667)):
            engine.set_option("amg_sweeps", sweeps)
            engine.set_option("amg_omega_permille", omega)
            w, st = engine.ot_solve(p.nu, stop="maxrel", tol=1e-9)
            assert st.status == "MA_OK" and st.max_rel_err <= 1e-9
            ws.append(w - w[-1])
    finally:
        engine.set_option("amg_sweeps", 2)
        engine.set_option("amg_omega_permille", 850)
    # the mass tolerance 1e-9 leaves the weights free at the level of the
    # Laplacian's conditioning times 1e-9
    for w in ws[1:]:
        assert np.abs(w - ws[0]).max() <= 1e-4 * np.abs(ws[0]).max()
    for name, bad in (("amg_sweeps", 3), ("amg_omega_permille", 50), ("amg_corr_permille", 4000)):
        with pytest.raises(EngineError):
            engine.set_option(name, bad)
