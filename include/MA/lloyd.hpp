// Drop-in for first_moment / second_moment / lloyd of include/MA/lloyd.hpp (MongeAmpere++).
#ifndef MA_LLOYD_HPP
#define MA_LLOYD_HPP

#include "engine_shim.hpp"

namespace MA {

// masses <- cell masses, centroids (N x 2) <- first moments int rho x, int rho y
// (include/MA/lloyd.hpp:30-69).
template <class T, class Functions, class Matrix, class Vector>
void first_moment(const T &densityT, const Functions &densityF, const Matrix &X,
                  const Vector &weights, Vector &masses, Matrix &centroids) {
  ma_engine *e = shim::engine();
  shim::set_mesh(e, densityT, densityF);
  const int64_t n = shim::set_diracs(e, X);
  const std::vector<double> w = shim::to_std(weights, n);
  std::vector<double> m(n), mom(2 * n);
  shim::check(e, ma_first_moment(e, w.data(), m.data(), mom.data()));
  shim::from_std(m, masses);
  centroids = Matrix::Zero(n, 2);
  for (int64_t i = 0; i < n; i++) {
    centroids(i, 0) = mom[2 * i];
    centroids(i, 1) = mom[2 * i + 1];
  }
}

// masses, centroids (N x 2) <- int rho x, int rho y, inertia (N x 3) <-
// int rho x^2, int rho y^2, int rho x y (include/MA/lloyd.hpp:71-123).
template <class T, class Functions, class Matrix, class Vector>
void second_moment(const T &densityT, const Functions &densityF, const Matrix &X,
                   const Vector &weights, Vector &masses, Matrix &centroids, Matrix &inertia) {
  ma_engine *e = shim::engine();
  shim::set_mesh(e, densityT, densityF);
  const int64_t n = shim::set_diracs(e, X);
  const std::vector<double> w = shim::to_std(weights, n);
  std::vector<double> m(n), mom(2 * n), mom2(3 * n);
  shim::check(e, ma_second_moment(e, w.data(), m.data(), mom.data(), mom2.data()));
  shim::from_std(m, masses);
  centroids = Matrix::Zero(n, 2);
  inertia = Matrix::Zero(n, 3);
  for (int64_t i = 0; i < n; i++) {
    centroids(i, 0) = mom[2 * i];
    centroids(i, 1) = mom[2 * i + 1];
    for (int c = 0; c < 3; c++) inertia(i, c) = mom2[3 * i + c];
  }
}

// Centroids of the cells: first moments divided by the masses
// (include/MA/lloyd.hpp:126-144).
template <class T, class Functions, class Matrix, class Vector>
void lloyd(const T &densityT, const Functions &densityF, const Matrix &X, const Vector &weights,
           Vector &masses, Matrix &centroids) {
  first_moment(densityT, densityF, X, weights, masses, centroids);
  const int64_t n = (int64_t)X.rows();
  for (int64_t i = 0; i < n; i++) {
    centroids(i, 0) /= masses(i);
    centroids(i, 1) /= masses(i);
  }
}

}  // namespace MA

#endif
