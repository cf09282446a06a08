// Drop-in for the polygon-materialising traversal of
// include/MA/voronoi_triangulation_intersection.hpp:315-343 (MongeAmpere++).
#ifndef MA_VORONOI_TRIANGULATION_INTERSECTION_HPP
#define MA_VORONOI_TRIANGULATION_INTERSECTION_HPP

#include <type_traits>
#include <vector>

#include "engine_shim.hpp"

namespace MA {

// Calls out(polygon, face, i) once per fragment of positive area: the
// Laguerre cell of Dirac i (X row i, weight weights(i)) cut by the face, in
// order of (i, face).  polygon is a std::vector of the triangulation's point
// type, counter-clockwise.  The reference's traversal takes the regular
// triangulation dt of the weighted Diracs and hands out its vertex handle
// (whose info() is the Dirac index); the engine builds the power diagram
// itself, so the Diracs and weights come in as for MA::kantorovich and the
// callback receives the index.
template <class T, class Functions, class Matrix, class Vector, class F>
void voronoi_triangulation_intersection(const T &densityT, const Functions &densityF,
                                        const Matrix &X, const Vector &weights, F out) {
  using Point =
      typename std::decay<decltype(densityT.finite_faces_begin()->vertex(0)->point())>::type;
  ma_engine *e = shim::engine();
  shim::set_mesh(e, densityT, densityF);
  const int64_t n = shim::set_diracs(e, X);
  const std::vector<double> w = shim::to_std(weights, n);
  int64_t nf = 0, nv = 0;
  const ma_status s = ma_fragments(e, w.data(), 0, 0, &nf, &nv, nullptr, nullptr, nullptr,
                                   nullptr);
  if (s != MA_CAPACITY && s != MA_OK) shim::check(e, s);
  std::vector<int32_t> cell(nf > 0 ? nf : 1), tri(nf > 0 ? nf : 1);
  std::vector<int64_t> start(nf + 1);
  std::vector<double> xy(nv > 0 ? 2 * nv : 2);
  shim::check(e, ma_fragments(e, w.data(), nf, nv, &nf, &nv, cell.data(), tri.data(),
                              start.data(), xy.data()));
  // face handles in the order the mesh was built from
  std::vector<typename std::decay<decltype(densityT.finite_faces_begin())>::type> faces;
  for (auto f = densityT.finite_faces_begin(); f != densityT.finite_faces_end(); ++f)
    faces.push_back(f);
  std::vector<Point> poly;
  for (int64_t f = 0; f < nf; f++) {
    poly.clear();
    for (int64_t q = start[f]; q < start[f + 1]; q++) poly.push_back(Point{xy[2 * q], xy[2 * q + 1]});
    out(poly, faces[tri[f]], (size_t)cell[f]);
  }
}

}  // namespace MA

#endif
