// Drives the header-only C++ shim (include/MA/*.hpp) with minimal stand-ins
// for the CGAL triangulation / linear-function map and the Eigen matrix and
// vector types that MongeAmpere++ instantiates its templates with.
//
// Input (text file): nv, then nv lines "x y value"; nt, then nt lines
// "i j k"; n, then n lines "x y nu".  Output: one JSON object with the
// results of MA::kantorovich (w = 0), MA::ot_solve, MA::solve_laplacian_matrix
// MA::lloyd, MA::second_moment, MA::voronoi_triangulation_intersection
// (fragment count and area per Dirac) and MA::draw_laguerre_diagram (a 16 x 12
// image of the unit square, colour of Dirac i = i + 1).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <vector>

#include "MA/kantorovich.hpp"
#include "MA/lloyd.hpp"
#include "MA/optimal_transport.hpp"
#include "MA/rasterization.hpp"
#include "MA/voronoi_triangulation_intersection.hpp"

namespace mock {

struct Point {
  double px, py;
  double x() const { return px; }
  double y() const { return py; }
};
struct Vertex {
  Point p;
  const Point &point() const { return p; }
};
using Vertex_handle = const Vertex *;
struct Face {
  Vertex_handle v[3];
  Vertex_handle vertex(int i) const { return v[i]; }
};
using Face_handle = const Face *;

struct Triangulation {
  std::vector<Vertex> verts;
  std::vector<Face> faces;
  struct Finite_faces_iterator {
    const Face *f;
    Face_handle operator->() const { return f; }
    operator Face_handle() const { return f; }
    Finite_faces_iterator &operator++() {
      ++f;
      return *this;
    }
    bool operator!=(const Finite_faces_iterator &o) const { return f != o.f; }
  };
  Finite_faces_iterator finite_faces_begin() const { return {faces.data()}; }
  Finite_faces_iterator finite_faces_end() const { return {faces.data() + faces.size()}; }
};

// density a x + b y + c on one face
struct Linear_function {
  double a = 0, b = 0, c = 0;
  double operator()(const Point &p) const { return a * p.x() + b * p.y() + c; }
};

struct Vector {
  std::vector<double> d;
  Vector() = default;
  explicit Vector(int64_t n) : d(n, 0.0) {}
  static Vector Zero(int64_t n) { return Vector(n); }
  int64_t size() const { return (int64_t)d.size(); }
  int64_t rows() const { return size(); }
  double &operator()(int64_t i) { return d[i]; }
  double operator()(int64_t i) const { return d[i]; }
};

struct Matrix {
  int64_t r = 0, c = 0;
  std::vector<double> d;  // column-major, like Eigen
  Matrix() = default;
  Matrix(int64_t r_, int64_t c_) : r(r_), c(c_), d(r_ * c_, 0.0) {}
  static Matrix Zero(int64_t r_, int64_t c_) { return Matrix(r_, c_); }
  int64_t rows() const { return r; }
  double &operator()(int64_t i, int64_t j) { return d[j * r + i]; }
  double operator()(int64_t i, int64_t j) const { return d[j * r + i]; }
};

// column-compressed, duplicates summed by setFromTriplets
struct SparseMatrix {
  int64_t n = 0;
  std::vector<std::map<int, double>> cols;
  SparseMatrix() = default;
  SparseMatrix(int64_t r, int64_t c) : n(r), cols(c) {}
  int64_t rows() const { return n; }
  template <class It>
  void setFromTriplets(It b, It e) {
    for (auto &c : cols) c.clear();
    for (; b != e; ++b) cols[b->col()][b->row()] += b->value();
  }
  struct InnerIterator {
    std::map<int, double>::const_iterator it, end;
    InnerIterator(const SparseMatrix &m, int64_t j) : it(m.cols[j].begin()), end(m.cols[j].end()) {}
    explicit operator bool() const { return it != end; }
    InnerIterator &operator++() {
      ++it;
      return *this;
    }
    int row() const { return it->first; }
    double value() const { return it->second; }
  };
};

}  // namespace mock

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s input.txt\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1]);
  mock::Triangulation T;
  int64_t nv, nt, n;
  in >> nv;
  std::vector<double> val(nv);
  T.verts.resize(nv);
  for (int64_t i = 0; i < nv; i++) in >> T.verts[i].p.px >> T.verts[i].p.py >> val[i];
  in >> nt;
  T.faces.resize(nt);
  std::map<mock::Face_handle, mock::Linear_function> F;
  for (int64_t f = 0; f < nt; f++) {
    int a, b, c;
    in >> a >> b >> c;
    T.faces[f].v[0] = &T.verts[a];
    T.faces[f].v[1] = &T.verts[b];
    T.faces[f].v[2] = &T.verts[c];
  }
  for (int64_t f = 0; f < nt; f++) {
    // plane through the three vertex values
    const mock::Face &fc = T.faces[f];
    const int ia = (int)(fc.v[0] - T.verts.data()), ib = (int)(fc.v[1] - T.verts.data()),
              ic = (int)(fc.v[2] - T.verts.data());
    const double x0 = fc.v[0]->p.px, y0 = fc.v[0]->p.py, x1 = fc.v[1]->p.px, y1 = fc.v[1]->p.py;
    const double x2 = fc.v[2]->p.px, y2 = fc.v[2]->p.py;
    const double D = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    const double f1 = val[ib] - val[ia], f2 = val[ic] - val[ia];
    mock::Linear_function L;
    L.a = (f1 * (y2 - y0) - f2 * (y1 - y0)) / D;
    L.b = ((x1 - x0) * f2 - (x2 - x0) * f1) / D;
    L.c = val[ia] - L.a * x0 - L.b * y0;
    F[&T.faces[f]] = L;
  }
  in >> n;
  mock::Matrix X(n, 2);
  mock::Vector nu(n);
  for (int64_t i = 0; i < n; i++) in >> X(i, 0) >> X(i, 1) >> nu(i);

  mock::Vector w0 = mock::Vector::Zero(n), g;
  mock::SparseMatrix h;
  const double fval = MA::kantorovich(T, F, X, w0, g, h);

  mock::Vector rhs(n);
  for (int64_t i = 0; i < n; i++) rhs(i) = g(i) - nu(i);
  const mock::Vector d = MA::solve_laplacian_matrix(h, rhs);

  mock::Vector x;
  MA::Statistics st{0, 0};
  MA::ot_solve(T, F, X, nu, x, 1e-10, 100, false, &st);

  mock::Vector m;
  mock::Matrix cen;
  MA::lloyd(T, F, X, x, m, cen);
  mock::Vector m2;
  mock::Matrix mom, inertia;
  MA::second_moment(T, F, X, x, m2, mom, inertia);

  std::vector<double> farea(n, 0.0);
  int64_t nfrag = 0;
  MA::voronoi_triangulation_intersection(
      T, F, X, x, [&](const std::vector<mock::Point> &poly, mock::Face_handle f, size_t i) {
        (void)f;
        double a = 0.0;
        for (size_t k = 0; k < poly.size(); k++) {
          const mock::Point &p = poly[k], &q = poly[(k + 1) % poly.size()];
          a += p.x() * q.y() - q.x() * p.y();
        }
        farea[i] += 0.5 * a;
        nfrag++;
      });
  std::vector<double> colors(n);
  for (int64_t i = 0; i < n; i++) colors[i] = (double)(i + 1);
  const int IW = 16, IH = 12;
  std::vector<double> image(IW * IH, 0.0);
  MA::draw_laguerre_diagram(T, F, X, x, colors, 0.0, 0.0, 1.0, 1.0, IW, IH,
                            [&](int px, int py, double v) { image[py * IW + px] += v; });

  std::printf("{\"fval\": %.17g, \"niter\": %zu, \"neval\": %zu,\n", fval, st.niter, st.neval);
  std::printf("\"nfrag\": %lld,\n", (long long)nfrag);
  std::printf("\"frag_area\": [");
  for (int64_t i = 0; i < n; i++) std::printf("%s%.17g", i ? ", " : "", farea[i]);
  std::printf("],\n\"image\": [");
  for (int k = 0; k < IW * IH; k++) std::printf("%s%.17g", k ? ", " : "", image[k]);
  std::printf("],\n");
  auto vec = [&](const char *name, const mock::Vector &v, bool last = false) {
    std::printf("\"%s\": [", name);
    for (int64_t i = 0; i < v.size(); i++) std::printf("%s%.17g", i ? ", " : "", v(i));
    std::printf("]%s\n", last ? "" : ",");
  };
  vec("g", g);
  vec("d", d);
  vec("w", x);
  vec("m", m);
  std::printf("\"centroids\": [");
  for (int64_t i = 0; i < n; i++)
    std::printf("%s[%.17g, %.17g]", i ? ", " : "", cen(i, 0), cen(i, 1));
  std::printf("],\n\"inertia\": [");
  for (int64_t i = 0; i < n; i++)
    std::printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", inertia(i, 0), inertia(i, 1),
                inertia(i, 2));
  std::printf("],\n\"hessian\": [");
  bool first = true;
  for (int64_t j = 0; j < n; j++)
    for (mock::SparseMatrix::InnerIterator it(h, j); it; ++it) {
      std::printf("%s[%d, %lld, %.17g]", first ? "" : ", ", it.row(), (long long)j, it.value());
      first = false;
    }
  std::printf("]}\n");
  return 0;
}
