// Drop-in for include/MA/optimal_transport.hpp of MongeAmpere++.
#ifndef MA_OPTIMAL_TRANSPORT_HPP
#define MA_OPTIMAL_TRANSPORT_HPP

#include <iostream>

#include "engine_shim.hpp"

namespace MA {

// include/MA/optimal_transport.hpp:35-39
struct Statistics {
  size_t niter;
  size_t neval;
};

// Gauge-fixed Newton direction: d[:N-1] solves h[:N-1,:N-1] d = g[:N-1],
// d[N-1] = 0 (include/MA/optimal_transport.hpp:41-87), by multigrid-
// preconditioned CG on the GPU.  h: any sparse type with outerSize()-free
// access through rows() and a row/col/value iteration (Eigen: InnerIterator).
template <class SparseMatrix, class Vector>
Vector solve_laplacian_matrix(const SparseMatrix &h, const Vector &g, bool verbose = false) {
  ma_engine *e = shim::engine();
  const int64_t n = (int64_t)h.rows();
  // CSR from the column-compressed symmetric matrix: column j of h is row j
  std::vector<int32_t> ptr(n + 1, 0), col;
  std::vector<double> val;
  for (int64_t j = 0; j < n; j++) {
    for (typename SparseMatrix::InnerIterator it(h, j); it; ++it) {
      col.push_back((int32_t)it.row());
      val.push_back(it.value());
    }
    ptr[j + 1] = (int32_t)col.size();
  }
  const std::vector<double> gs = shim::to_std(g, n);
  std::vector<double> d(n);
  double res = 0.0;
  int64_t iters = 0;
  shim::check(e, ma_solve_laplacian_csr(e, n, ptr.data(), col.data(), val.data(), gs.data(),
                                        d.data(), 1e-12, 0, &res, &iters));
  if (verbose) std::cerr << "CG: " << iters << " iterations, relative residual " << res << "\n";
  Vector out;
  shim::from_std(d, out);
  return out;
}

// Damped Newton solve of the semi-discrete transport problem
// (include/MA/optimal_transport.hpp:89-193): x holds the initial guess (used
// when its size is N) and receives the weights; stops when ||m - nu||_2 <
// eps_g or after maxiter + 1 steps; an empty initial cell prints the
// reference's message and returns with x unchanged.
template <class T, class Functions, class Matrix, class Vector>
void ot_solve(const T &t, const Functions &functions, const Matrix &X, const Vector &masses,
              Vector &x, double eps_g = 1e-7, size_t maxiter = 100, bool verbose = true,
              struct Statistics *stats = 0) {
  ma_engine *e = shim::engine();
  shim::set_mesh(e, t, functions);
  const int64_t n = shim::set_diracs(e, X);
  const bool init = (int64_t)x.size() == n;
  std::vector<double> w = init ? shim::to_std(x, n) : std::vector<double>(n, 0.0);
  const std::vector<double> nu = shim::to_std(masses, n);
  ma_solve_opts o;
  ma_default_solve_opts(&o);
  o.precision = shim::precision();
  o.stop_rule = MA_STOP_L2;
  o.eps_g = eps_g;
  o.maxiter = (int)maxiter;
  o.verbose = verbose ? 1 : 0;
  ma_stats st;
  const ma_status s = ma_ot_solve(e, nu.data(), w.data(), init ? 1 : 0, &o, &st);
  if (s == MA_EMPTY_CELL) {
    std::cerr << "Error: computed minimum mass is non-positive\n"
              << "This is because the Laguerre cell for the point [" << X(st.bad_index, 0)
              << ", " << X(st.bad_index, 1) << "], i=" << st.bad_index << " is empty.\n";
    if (!init) x = Vector::Zero(n);
    return;
  }
  if (s != MA_OK && s != MA_NOT_CONVERGED) shim::check(e, s);
  shim::from_std(w, x);
  if (stats) {
    stats->niter = (size_t)st.niter;
    stats->neval = (size_t)st.neval;
  }
}

}  // namespace MA

#endif
