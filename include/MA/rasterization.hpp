// Drop-in for draw_laguerre_diagram of include/MA/rasterization.hpp:491-547
// (MongeAmpere++).
#ifndef MA_RASTERIZATION_HPP
#define MA_RASTERIZATION_HPP

#include <type_traits>
#include <vector>

#include "engine_shim.hpp"

namespace MA {

namespace shim {
template <class C>
int channels(const C &c, std::true_type) {
  (void)c;
  return 1;
}
template <class C>
int channels(const C &c, std::false_type) {
  return (int)c.size();
}
template <class C>
double channel(const C &c, int k, std::true_type) {
  (void)k;
  return (double)c;
}
template <class C>
double channel(const C &c, int k, std::false_type) {
  return (double)c[k];
}
template <class C>
void set_channel(C &c, int k, double v, std::true_type) {
  (void)k;
  c = (C)v;
}
template <class C>
void set_channel(C &c, int k, double v, std::false_type) {
  c[k] = v;
}
}  // namespace shim

// The box [x0,x1] x [y0,y1] drawn into a w x h pixel grid: every fragment
// (Laguerre cell i cut by a face) contributes, to each pixel (x, y) it
// covers, cover x (mean density over its vertices) x colors[i], cover being
// the covered fraction of the pixel.  The reference hands each fragment's
// contribution to put_pixel(x, y, value) separately; here put_pixel is called
// once per touched pixel with the sum of the contributions, which is the same
// picture for the summing put_pixel the reference's callers use.  A colour is
// a number or a fixed-size vector (size() and operator[]).
template <class T, class Functions, class Matrix, class Vector, class ColorVector,
          class PutPixel>
void draw_laguerre_diagram(const T &densityT, const Functions &densityF, const Matrix &X,
                           const Vector &weights, const ColorVector &colors, double x0, double y0,
                           double x1, double y1, size_t w, size_t h, PutPixel put_pixel) {
  using Color = typename std::decay<decltype(colors[0])>::type;
  using Scalar = typename std::is_arithmetic<Color>::type;
  ma_engine *e = shim::engine();
  shim::set_mesh(e, densityT, densityF);
  const int64_t n = shim::set_diracs(e, X);
  const std::vector<double> wt = shim::to_std(weights, n);
  const int nc = shim::channels(colors[0], Scalar());
  std::vector<double> col((size_t)n * nc);
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < nc; k++) col[i * nc + k] = shim::channel(colors[i], k, Scalar());
  std::vector<double> img(w * h * nc);
  shim::check(e, ma_draw_laguerre(e, wt.data(), col.data(), nc, x0, y0, x1, y1, (int)w, (int)h,
                                  img.data(), 0));
  Color c = colors[0];
  for (size_t y = 0; y < h; y++)
    for (size_t x = 0; x < w; x++) {
      const double *p = &img[(y * w + x) * nc];
      bool any = false;
      for (int k = 0; k < nc; k++) any |= p[k] != 0.0;
      if (!any) continue;
      for (int k = 0; k < nc; k++) shim::set_channel(c, k, p[k], Scalar());
      put_pixel((int)x, (int)y, c);
    }
}

}  // namespace MA

#endif
