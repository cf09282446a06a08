// Shared plumbing of the header-only C++ shim that keeps the MongeAmpere++
// call shapes (MA::kantorovich, MA::ot_solve, MA::solve_laplacian_matrix,
// MA::first_moment, MA::lloyd) on top of the C ABI of include/ma/ma.h.
//
// The shim is generic over the reference's template parameters:
//   T          a triangulation with finite_faces_begin()/finite_faces_end();
//              a face handle f gives f->vertex(i) (i = 0..2), a vertex handle
//              gives point() with x() and y(); vertex handles are ordered
//              (operator<), as CGAL handles are;
//   Functions  an associative container face handle -> callable density,
//              find(f) / end() (the std::map<Face_handle, Linear_function> of
//              include/MA/functions.hpp:55-80);
//   Matrix     rows(), operator()(i, j) (Eigen::MatrixXd);
//   Vector     size()/rows(), operator()(i) / operator[](i), Vector::Zero(n);
//   SparseMatrix  SparseMatrix(n, n) and setFromTriplets(begin, end) over
//              objects with row(), col(), value() (Eigen::SparseMatrix<double>).
//
// One engine per thread, on device MA_DEVICE (environment, default 0).  The
// precision of MA::kantorovich and MA::ot_solve is FP64 unless the
// environment sets MA_PRECISION=mixed.  Errors throw std::runtime_error,
// except where the reference itself prints and returns (ot_solve).
#ifndef MA_ENGINE_SHIM_HPP
#define MA_ENGINE_SHIM_HPP

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../ma/ma.h"

namespace MA {
namespace shim {

inline ma_engine *engine() {
  struct Holder {
    ma_engine *e = nullptr;
    ~Holder() {
      if (e) ma_destroy(e);
    }
  };
  thread_local Holder h;
  if (!h.e) {
    const char *dev = std::getenv("MA_DEVICE");
    if (ma_create(dev ? std::atoi(dev) : 0, &h.e) != MA_OK)
      throw std::runtime_error("MA engine: no usable GPU");
  }
  return h.e;
}

inline int precision() {
  const char *p = std::getenv("MA_PRECISION");
  return (p && std::strcmp(p, "mixed") == 0) ? MA_MIXED : MA_FP64;
}

inline void check(ma_engine *e, ma_status s) {
  if (s != MA_OK) throw std::runtime_error(std::string("MA engine: ") + ma_last_error(e));
}

struct Triplet {
  int r, c;
  double v;
  int row() const { return r; }
  int col() const { return c; }
  double value() const { return v; }
};

// Mesh with P1 values from (T, Functions).  Each face's density is evaluated
// at its own vertices; when every face agrees at every shared vertex (a
// continuous P1 density) the vertices are shared, otherwise each face gets
// its own three vertices, which reproduces a discontinuous piecewise-linear
// density exactly.
template <class T, class Functions>
void set_mesh(ma_engine *e, const T &t, const Functions &F) {
  using VH = typename std::decay<decltype(t.finite_faces_begin()->vertex(0))>::type;
  std::map<VH, int32_t> id;
  std::vector<double> vxy, vval, fxy, fval;
  std::vector<int32_t> tri;
  bool continuous = true;
  for (auto f = t.finite_faces_begin(); f != t.finite_faces_end(); ++f) {
    auto fit = F.find(f);
    if (fit == F.end()) throw std::runtime_error("MA shim: face without density function");
    for (int i = 0; i < 3; i++) {
      auto v = f->vertex(i);
      const auto &p = v->point();
      const double x = p.x(), y = p.y(), val = fit->second(p);
      fxy.push_back(x);
      fxy.push_back(y);
      fval.push_back(val);
      auto it = id.find(v);
      if (it == id.end()) {
        const int32_t k = (int32_t)vval.size();
        id.emplace(v, k);
        vxy.push_back(x);
        vxy.push_back(y);
        vval.push_back(val);
        tri.push_back(k);
      } else {
        const double u = vval[it->second];
        if (std::fabs(u - val) > 1e-12 * (std::fabs(u) + std::fabs(val))) continuous = false;
        tri.push_back(it->second);
      }
    }
  }
  const int64_t nt = (int64_t)tri.size() / 3;
  if (continuous) {
    check(e, ma_set_mesh(e, (int64_t)vval.size(), vxy.data(), vval.data(), nt, tri.data()));
  } else {
    std::vector<int32_t> ftri(tri.size());
    for (size_t k = 0; k < ftri.size(); k++) ftri[k] = (int32_t)k;
    check(e, ma_set_mesh(e, (int64_t)fval.size(), fxy.data(), fval.data(), nt, ftri.data()));
  }
}

template <class Matrix>
int64_t set_diracs(ma_engine *e, const Matrix &X) {
  const int64_t n = (int64_t)X.rows();
  std::vector<double> xy(2 * n);
  for (int64_t i = 0; i < n; i++) {
    xy[2 * i] = X(i, 0);
    xy[2 * i + 1] = X(i, 1);
  }
  check(e, ma_set_diracs(e, n, xy.data(), MA_XY_INTERLEAVED));
  return n;
}

template <class Vector>
std::vector<double> to_std(const Vector &v, int64_t n) {
  std::vector<double> o(n);
  for (int64_t i = 0; i < n; i++) o[i] = v(i);
  return o;
}

template <class Vector>
void from_std(const std::vector<double> &s, Vector &v) {
  const int64_t n = (int64_t)s.size();
  v = Vector::Zero(n);
  for (int64_t i = 0; i < n; i++) v(i) = s[i];
}

// The engine's last Hessian as a SparseMatrix (caller order).
template <class SparseMatrix>
void hessian(ma_engine *e, int64_t n, SparseMatrix &h) {
  ma_eval_info info;
  check(e, ma_get_eval_info(e, &info));
  std::vector<int32_t> ptr(n + 1), col(info.nnz > 0 ? info.nnz : 1);
  std::vector<double> val(col.size());
  check(e, ma_hessian_copy(e, ptr.data(), col.data(), val.data()));
  std::vector<Triplet> t;
  t.reserve((size_t)info.nnz);
  for (int64_t i = 0; i < n; i++)
    for (int32_t q = ptr[i]; q < ptr[i + 1]; q++) t.push_back({(int)i, col[q], val[q]});
  h = SparseMatrix(n, n);
  h.setFromTriplets(t.begin(), t.end());
}

}  // namespace shim
}  // namespace MA

#endif
