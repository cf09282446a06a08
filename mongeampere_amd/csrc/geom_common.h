// Device helpers shared by the geometry translation units.
#pragma once

#include <cmath>
#include <cstdint>

#include "engine.h"

namespace ma {

constexpr int kBlock = 128;
constexpr double kInf = HUGE_VAL;
// relative tolerance under which a point counts as lying on a triangle edge
constexpr double kSnap = 1e-13;

__host__ __device__ inline uint32_t part1by1(uint32_t x) {
  x &= 0x0000ffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__host__ __device__ inline uint32_t compact1by1(uint32_t x) {
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0f0f0f0fu;
  x = (x | (x >> 4)) & 0x00ff00ffu;
  x = (x | (x >> 8)) & 0x0000ffffu;
  return x;
}
__host__ __device__ inline uint32_t morton2(uint32_t x, uint32_t y) {
  return part1by1(x) | (part1by1(y) << 1);
}
__device__ inline int bin_coord(double v, double o, double h, int G) {
  double f = floor((v - o) / h);
  return (f < 0.0) ? 0 : (f > (double)(G - 1) ? G - 1 : (int)f);
}
static inline unsigned nblocks(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// Sparse row times vector by a group of kRowLanes consecutive lanes: lane l
// takes entries l, l + kRowLanes, ... of row i, so a wave reads consecutive
// entries of consecutive rows (coalesced); every lane of the group returns
// the sum.  All lanes of the group must call it (valid = false: no row).
constexpr int kRowLanes = 8;
__device__ inline double row_dot(const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                                 const double *__restrict__ val, const double *__restrict__ x,
                                 int64_t i, bool valid, int lane) {
  double s = 0.0;
  if (valid)
    for (int32_t q = ptr[i] + lane; q < ptr[i + 1]; q += kRowLanes) s += val[q] * x[col[q]];
  for (int o = 1; o < kRowLanes; o <<= 1) s += __shfl_xor(s, o, kRowLanes);
  return s;
}

__device__ inline void tri_bbox(const TriData &t, double &x0, double &y0, double &x1,
                                double &y1) {
  x0 = fmin(t.x[0], fmin(t.x[1], t.x[2]));
  x1 = fmax(t.x[0], fmax(t.x[1], t.x[2]));
  y0 = fmin(t.y[0], fmin(t.y[1], t.y[2]));
  y1 = fmax(t.y[0], fmax(t.y[1], t.y[2]));
}

// Offsets of the levels of the max-weight pyramid (level 0 = bins).
struct PyrOff {
  int64_t o[16];
};

// Albrecht-Collatz 6-point rule points (include/MA/quadrature.hpp:23-42):
// weights 1/30 for points 0..2, 9/30 for points 3..5, times the area.
template <class R>
struct AC6T {
  R px[6], py[6];
};
using AC6 = AC6T<double>;
template <class R>
__device__ inline void ac6_points(R ax, R ay, R bx, R by, R cx, R cy, AC6T<R> &q) {
  const R ux = bx - ax, uy = by - ay, vx = cx - ax, vy = cy - ay;
  const R h = R(0.5), s6 = R(1.0 / 6.0), t3 = R(2.0 / 3.0);
  q.px[0] = ax + h * ux + h * vx;
  q.py[0] = ay + h * uy + h * vy;
  q.px[1] = ax + h * ux;
  q.py[1] = ay + h * uy;
  q.px[2] = ax + h * vx;
  q.py[2] = ay + h * vy;
  q.px[3] = ax + s6 * ux + t3 * vx;
  q.py[3] = ay + s6 * uy + t3 * vy;
  q.px[4] = ax + s6 * vx + t3 * ux;
  q.py[4] = ay + s6 * vy + t3 * uy;
  q.px[5] = ax + s6 * ux + s6 * vx;
  q.py[5] = ay + s6 * uy + s6 * vy;
}

// Enumerates, once each, the triangles whose bounding box meets the box
// [px0,px1]x[py0,py1] (absolute), scanning the bins of that box clipped to
// the bin rectangle [rx0,rx1]x[ry0,ry1].  A triangle listed in several bins
// is visited in the bin holding the lower-left corner of the overlap of the
// two boxes.
// tbox: per triangle its bounding box (x0, y0, x1, y1) in FP32, rounded
// outwards (16 B per test instead of the 72 B triangle record).
template <class F>
__device__ inline void for_each_triangle(const Grid &g, const int32_t *__restrict__ tbin_start,
                                         const int32_t *__restrict__ tbin_items,
                                         const float4 *__restrict__ tbox, double px0, double py0,
                                         double px1, double py1, int rx0, int ry0, int rx1,
                                         int ry1, F f) {
  const int qx0 = max(bin_coord(px0, g.x0, g.h, g.G), rx0);
  const int qx1 = min(bin_coord(px1, g.x0, g.h, g.G), rx1);
  const int qy0 = max(bin_coord(py0, g.y0, g.h, g.G), ry0);
  const int qy1 = min(bin_coord(py1, g.y0, g.h, g.G), ry1);
  for (int by = qy0; by <= qy1; by++)
    for (int bx = qx0; bx <= qx1; bx++) {
      const int64_t b = morton2((uint32_t)bx, (uint32_t)by);
      for (int32_t q = tbin_start[b]; q < tbin_start[b + 1]; q++) {
        const int32_t ti = tbin_items[q];
        const float4 tb = tbox[ti];
        const double tx0 = tb.x, ty0 = tb.y, tx1 = tb.z, ty1 = tb.w;
        if (tx1 < px0 || tx0 > px1 || ty1 < py0 || ty0 > py1) continue;
        const int ox = max(bin_coord(fmax(tx0, px0), g.x0, g.h, g.G), qx0);
        const int oy = max(bin_coord(fmax(ty0, py0), g.y0, g.h, g.G), qy0);
        if (ox != bx || oy != by) continue;
        f(ti);
      }
    }
}

}  // namespace ma
