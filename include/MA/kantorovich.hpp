// Drop-in for include/MA/kantorovich.hpp of MongeAmpere++: same name, template
// parameters and argument meaning, evaluated by the MI355X engine.
#ifndef MA_KANTOROVICH_HPP
#define MA_KANTOROVICH_HPP

#include "engine_shim.hpp"

namespace MA {

// Value of the Kantorovich functional at weights w; g <- cell masses
// (reassigned, as include/MA/kantorovich.hpp:83), h <- the N x N Hessian
// H_ij = -int_{edge ij} rho / (2 |y_i - y_j|), H_ii = -sum_j H_ij
// (include/MA/kantorovich.hpp:35-141).
template <class T, class Functions, class Matrix, class Vector, class SparseMatrix>
double kantorovich(const T &densityT, const Functions &densityF, const Matrix &X,
                   const Vector &weights, Vector &g, SparseMatrix &h) {
  ma_engine *e = shim::engine();
  shim::set_mesh(e, densityT, densityF);
  const int64_t n = shim::set_diracs(e, X);
  const std::vector<double> w = shim::to_std(weights, n);
  std::vector<double> m(n);
  double fval = 0.0;
  int64_t nnz = 0;
  shim::check(e, ma_kantorovich(e, w.data(), shim::precision(), &fval, m.data(), &nnz));
  shim::from_std(m, g);
  shim::hessian(e, n, h);
  return fval;
}

}  // namespace MA

#endif
