// Register/LDS-resident evaluation kernels:
//   k_cell_build  one thread per Dirac: power cell clipped in place in LDS
//                 (FP64, at most kMaxFast vertices), written once to HBM;
//   k_pairs       (cell, triangle) pairs whose bounding boxes meet;
//   k_frag        one thread per pair: the cell polygon streamed through a
//                 three-stage Sutherland-Hodgman pipeline (one stage per
//                 triangle edge) held in registers, fused with the exact
//                 quadrature of the fragment;
//   k_rows        one thread per Dirac: Hessian row from the edge integrals.
// Cells that overflow kMaxFast vertices are flagged and evaluated by the
// general kernel (k_cells_general).
//
// Reference behaviour (paths relative to the MongeAmpere++ tree):
//   strict tie rule, edge labels   include/MA/voronoi_triangulation_intersection.hpp:125-198
//   per-fragment integrals         include/MA/kantorovich.hpp:109-135, include/MA/lloyd.hpp:45-60
//   quadrature                     include/MA/quadrature.hpp:23-42,69-85,122-143
#include <climits>

#include "geom_common.h"

namespace ma {

constexpr int kMaxFast = 16;
constexpr int kCellBlock = 128;
constexpr int kLocalRings = 3;  // site-centric search radius, in bins

struct CellArgs {
  Grid g;
  int64_t n;
  const double *pos, *w;
  const int32_t *perm, *rect, *cand_cnt, *cand;
  int cap;
  const int32_t *tbin_start, *tbin_items, *bin_start;
  double dom[4];     // domain bounding box
  const TriData *tri;
  double *px, *py;   // cell vertices (local to the Dirac), [v * n + k]
  int32_t *plab;     // neighbour across edge v -> v+1 (-1: bin-rectangle edge)
  int32_t *pn;       // vertex count, -1 = handled by the general kernel
  double *bbox;      // absolute bounding box of the cell (4 per Dirac)
  int32_t *tcount;   // candidate triangles per cell
  double *mass, *intg, *mom;  // mom: local first moments, interleaved (x, y)
  double *eacc;      // edge integrals int rho ds, [v * n + k]
  int32_t *ehit;     // edge seen in a fragment
  int32_t *flags;
  int force_general;  // route every cell to the general kernel (testing)
};

// In-place clip of the LDS polygon of thread t by {X : kk - 2 X.Y > 0}
// (strict: points with pi_k == pi_j are outside).  The surviving vertices
// form one cyclic run; it is moved to the front, followed by the exit point
// X (label = new neighbour) and the entry point E (label of the edge it lies
// on).  Returns the new vertex count; sets fail on overflow or when rounding
// breaks the single-run structure.
__device__ inline int clip_lds(double (*sx)[kCellBlock], double (*sy)[kCellBlock],
                               int (*sl)[kCellBlock], int t, int n, double Yx, double Yy,
                               double kk, int label, bool &fail) {
  uint32_t in = 0u;
  for (int v = 0; v < n; v++) {
    const double s = kk - 2.0 * (sx[v][t] * Yx + sy[v][t] * Yy);
    if (s > 0.0) in |= 1u << v;
  }
  const uint32_t full = (1u << n) - 1u;
  if (in == full) return n;
  if (in == 0u) return 0;
  const uint32_t nxt = ((in >> 1) | ((in & 1u) << (n - 1))) & full;
  const uint32_t exit_m = in & ~nxt, entry_m = ~in & nxt & full;
  if (__popc(exit_m) != 1 || __popc(entry_m) != 1) {
    fail = true;
    return n;
  }
  const int xk = __ffs(exit_m) - 1, ek = __ffs(entry_m) - 1;
  const int xk1 = (xk + 1 == n) ? 0 : xk + 1, ek1 = (ek + 1 == n) ? 0 : ek + 1;
  const int cnt_in = __popc(in);
  const int m = cnt_in + 2;
  if (m > kMaxFast) {
    fail = true;
    return n;
  }
  double ax = sx[xk][t], ay = sy[xk][t], bx = sx[xk1][t], by = sy[xk1][t];
  double sa = kk - 2.0 * (ax * Yx + ay * Yy), sb = kk - 2.0 * (bx * Yx + by * Yy);
  double tt = sa / (sa - sb);
  const double Xx = ax + tt * (bx - ax), Xy = ay + tt * (by - ay);
  ax = sx[ek][t]; ay = sy[ek][t]; bx = sx[ek1][t]; by = sy[ek1][t];
  sa = kk - 2.0 * (ax * Yx + ay * Yy);
  sb = kk - 2.0 * (bx * Yx + by * Yy);
  tt = sa / (sa - sb);
  const double Ex = ax + tt * (bx - ax), Ey = ay + tt * (by - ay);
  const int El = sl[ek][t];
  const int s0 = ek1;
  int px, pe;
  if (s0 <= xk) {
    for (int i = 0; i < cnt_in; i++) {
      sx[i][t] = sx[s0 + i][t];
      sy[i][t] = sy[s0 + i][t];
      sl[i][t] = sl[s0 + i][t];
    }
    px = cnt_in;
    pe = cnt_in + 1;
  } else {
    const int tl = n - s0, dest = xk + 3;
    if (dest < s0) {
      for (int i = 0; i < tl; i++) {
        sx[dest + i][t] = sx[s0 + i][t];
        sy[dest + i][t] = sy[s0 + i][t];
        sl[dest + i][t] = sl[s0 + i][t];
      }
    } else if (dest > s0) {
      for (int i = tl - 1; i >= 0; i--) {
        sx[dest + i][t] = sx[s0 + i][t];
        sy[dest + i][t] = sy[s0 + i][t];
        sl[dest + i][t] = sl[s0 + i][t];
      }
    }
    px = xk + 1;
    pe = xk + 2;
  }
  sx[px][t] = Xx;
  sy[px][t] = Xy;
  sl[px][t] = label;
  sx[pe][t] = Ex;
  sy[pe][t] = Ey;
  sl[pe][t] = El;
  return m;
}

// The LDS polygon of one thread.
struct CellLDS {
  double (*sx)[kCellBlock];
  double (*sy)[kCellBlock];
  int (*sl)[kCellBlock];
  int t;
  __device__ double rho2(int n) const {
    double r = 0.0;
    for (int v = 0; v < n; v++) r = fmax(r, sx[v][t] * sx[v][t] + sy[v][t] * sy[v][t]);
    return r;
  }
  __device__ void rect(double X0, double Y0, double X1, double Y1) {
    sx[0][t] = X0; sy[0][t] = Y0; sl[0][t] = -1;
    sx[1][t] = X1; sy[1][t] = Y0; sl[1][t] = -1;
    sx[2][t] = X1; sy[2][t] = Y1; sl[2][t] = -1;
    sx[3][t] = X0; sy[3][t] = Y1; sl[3][t] = -1;
  }
};

// Clips the cell of Dirac k by the power half-plane of candidate j.  A cheap
// distance test first discards candidates that cannot reach the polygon:
// pi_j - pi_k >= d^2 - 2 rho d + w_k - w_j on the disc of radius rho.
__device__ inline void consider(const CellArgs &a, int64_t k, int32_t j, double yx, double yy,
                                double wk, CellLDS &L, int &n, double &rho2, bool &fail) {
  if (j == k || n == 0) return;
  const double Yx = a.pos[2 * j] - yx, Yy = a.pos[2 * j + 1] - yy, wj = a.w[j];
  const double d2 = Yx * Yx + Yy * Yy;
  if (d2 == 0.0) {
    // coincident Diracs: the larger weight (then the smaller caller index)
    // hides the other
    if (wj > wk || (wj == wk && a.perm[j] < a.perm[k])) n = 0;
    return;
  }
  const double d = sqrt(d2), rho = sqrt(rho2);
  if (d2 - 2.0 * rho * d + wk - wj > 1e-12 * (d2 + fabs(wk) + fabs(wj))) return;
  for (int v = 0; v < n; v++)
    if (L.sl[v][L.t] == j) return;  // already an edge of the cell
  n = clip_lds(L.sx, L.sy, L.sl, L.t, n, Yx, Yy, d2 + wk - wj, j, fail);
  rho2 = L.rho2(n);
}

// Writes a finished cell: polygon, bounding box, candidate triangle count.
__device__ inline void write_cell(const CellArgs &a, int64_t k, double yx, double yy,
                                  const CellLDS &L, int n) {
  const int64_t N = a.n;
  a.pn[k] = n;
  if (n == 0) {
    a.tcount[k] = 0;
    return;
  }
  double bx0 = kInf, by0 = kInf, bx1 = -kInf, by1 = -kInf;
  for (int v = 0; v < n; v++) {
    const double x = L.sx[v][L.t], y = L.sy[v][L.t];
    a.px[v * N + k] = x;
    a.py[v * N + k] = y;
    a.plab[v * N + k] = L.sl[v][L.t];
    a.eacc[v * N + k] = 0.0;
    a.ehit[v * N + k] = 0;
    bx0 = fmin(bx0, x);
    bx1 = fmax(bx1, x);
    by0 = fmin(by0, y);
    by1 = fmax(by1, y);
  }
  bx0 += yx; bx1 += yx; by0 += yy; by1 += yy;
  a.bbox[4 * k] = bx0;
  a.bbox[4 * k + 1] = by0;
  a.bbox[4 * k + 2] = bx1;
  a.bbox[4 * k + 3] = by1;
  int cnt = 0;
  for_each_triangle(a.g, a.tbin_start, a.tbin_items, a.tri, bx0, by0, bx1, by1, 0, 0,
                    a.g.G - 1, a.g.G - 1, [&](int32_t) { cnt++; });
  a.tcount[k] = cnt;
}

__device__ inline void zero_cell_outputs(const CellArgs &a, int64_t k) {
  a.mass[k] = 0.0;
  a.intg[k] = 0.0;
  a.mom[2 * k] = 0.0;
  a.mom[2 * k + 1] = 0.0;
}

// Site-centric construction: the domain box is clipped by the Diracs of the
// bins in Chebyshev rings around the Dirac's own bin.  After each ring the
// cell is certified if every unvisited Dirac j (distance >= D, weight <= W)
// satisfies D^2 - 2 rho D - (W - w_k) > 0, i.e. cannot cut the polygon.
// Uncertified cells are marked pn = -2 for the bin-candidate construction.
__global__ void __launch_bounds__(kCellBlock) k_cell_local(CellArgs a, const double *wmax,
                                                           int rmax) {
  __shared__ double sx[kMaxFast][kCellBlock];
  __shared__ double sy[kMaxFast][kCellBlock];
  __shared__ int sl[kMaxFast][kCellBlock];
  CellLDS L{sx, sy, sl, (int)threadIdx.x};
  const int64_t k = blockIdx.x * (int64_t)kCellBlock + threadIdx.x;
  if (k >= a.n) return;
  const Grid &g = a.g;
  zero_cell_outputs(a, k);
  if (a.force_general) {
    a.pn[k] = -1;
    a.tcount[k] = 0;
    atomicOr(&a.flags[3], 1);
    return;
  }
  const double yx = a.pos[2 * k], yy = a.pos[2 * k + 1], wk = a.w[k];
  L.rect(a.dom[0] - yx, a.dom[1] - yy, a.dom[2] - yx, a.dom[3] - yy);
  int n = 4;
  bool fail = false, certified = false;
  double rho2 = L.rho2(4);
  const double W = *wmax;
  const int sbx = bin_coord(yx, g.x0, g.h, g.G), sby = bin_coord(yy, g.y0, g.h, g.G);
  for (int r = 0; r <= rmax && n > 0 && !fail; r++) {
    for (int by = max(sby - r, 0); by <= min(sby + r, g.G - 1) && n > 0 && !fail; by++) {
      const bool edge_row = (by == sby - r) || (by == sby + r);
      for (int bx = max(sbx - r, 0); bx <= min(sbx + r, g.G - 1) && n > 0 && !fail; bx++) {
        if (!edge_row && bx != sbx - r && bx != sbx + r) continue;
        const int64_t b = morton2((uint32_t)bx, (uint32_t)by);
        for (int32_t q = a.bin_start[b]; q < a.bin_start[b + 1] && n > 0 && !fail; q++)
          consider(a, k, q, yx, yy, wk, L, n, rho2, fail);
      }
    }
    if (n == 0 || fail) break;
    double D = kInf;
    if (sbx - r > 0) D = fmin(D, yx - (g.x0 + (sbx - r) * g.h));
    if (sbx + r < g.G - 1) D = fmin(D, g.x0 + (sbx + r + 1) * g.h - yx);
    if (sby - r > 0) D = fmin(D, yy - (g.y0 + (sby - r) * g.h));
    if (sby + r < g.G - 1) D = fmin(D, g.y0 + (sby + r + 1) * g.h - yy);
    const double rho = sqrt(rho2);
    if (D == kInf ||
        (D > rho && D * D - 2.0 * rho * D - (W - wk) > 1e-12 * (D * D + fabs(W) + fabs(wk)))) {
      certified = true;
      break;
    }
  }
  if (fail) {
    a.pn[k] = -1;
    a.tcount[k] = 0;
    atomicOr(&a.flags[3], 1);
    return;
  }
  if (!certified && n > 0) {
    a.pn[k] = -2;
    a.tcount[k] = 0;
    atomicOr(&a.flags[2], 1);
    return;
  }
  write_cell(a, k, yx, yy, L, n);
}

// Bin-candidate construction for the cells listed in `list` (those the
// site-centric search could not certify): the bin rectangle of the Dirac is
// clipped by the candidate owners of its bins, nearest rings first.  The
// rectangle covers every bin where the Dirac may own a point and the owners
// of every point of these bins are among the candidates, so no certificate
// is needed.
__global__ void __launch_bounds__(kCellBlock) k_cell_bins(CellArgs a, const int32_t *list,
                                                          const int32_t *nlist) {
  __shared__ double sx[kMaxFast][kCellBlock];
  __shared__ double sy[kMaxFast][kCellBlock];
  __shared__ int sl[kMaxFast][kCellBlock];
  CellLDS L{sx, sy, sl, (int)threadIdx.x};
  const int64_t idx = blockIdx.x * (int64_t)kCellBlock + threadIdx.x;
  if (idx >= *nlist) return;
  const int64_t k = list[idx];
  const Grid &g = a.g;
  const int rx0 = a.rect[4 * k], ry0 = a.rect[4 * k + 1], rx1 = a.rect[4 * k + 2],
            ry1 = a.rect[4 * k + 3];
  if (rx0 > rx1) {
    a.pn[k] = 0;
    a.tcount[k] = 0;
    return;
  }
  const double yx = a.pos[2 * k], yy = a.pos[2 * k + 1], wk = a.w[k];
  L.rect(g.x0 + rx0 * g.h - yx, g.y0 + ry0 * g.h - yy, g.x0 + (rx1 + 1) * g.h - yx,
         g.y0 + (ry1 + 1) * g.h - yy);
  int n = 4;
  bool fail = false;
  double rho2 = L.rho2(4);
  const int sbx = min(max(bin_coord(yx, g.x0, g.h, g.G), rx0), rx1);
  const int sby = min(max(bin_coord(yy, g.y0, g.h, g.G), ry0), ry1);
  const int maxr = max(max(sbx - rx0, rx1 - sbx), max(sby - ry0, ry1 - sby));
  for (int r = 0; r <= maxr && n > 0 && !fail; r++) {
    for (int by = max(sby - r, ry0); by <= min(sby + r, ry1) && n > 0 && !fail; by++) {
      const bool edge_row = (by == sby - r) || (by == sby + r);
      for (int bx = max(sbx - r, rx0); bx <= min(sbx + r, rx1) && n > 0 && !fail; bx++) {
        if (!edge_row && bx != sbx - r && bx != sbx + r) continue;
        const int64_t b = morton2((uint32_t)bx, (uint32_t)by);
        const int c = a.cand_cnt[b];
        const int32_t *cl = a.cand + b * (int64_t)a.cap;
        for (int i = 0; i < c && n > 0 && !fail; i++)
          consider(a, k, cl[i], yx, yy, wk, L, n, rho2, fail);
      }
    }
  }
  if (fail) {
    a.pn[k] = -1;
    a.tcount[k] = 0;
    atomicOr(&a.flags[3], 1);
    return;
  }
  write_cell(a, k, yx, yy, L, n);
}

__global__ void k_pairs(CellArgs a, const int32_t *poff, int32_t *pair_cell, int32_t *pair_tri) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  if (a.tcount[k] == 0) return;
  int32_t o = poff[k];
  for_each_triangle(a.g, a.tbin_start, a.tbin_items, a.tri, a.bbox[4 * k], a.bbox[4 * k + 1],
                    a.bbox[4 * k + 2], a.bbox[4 * k + 3], 0, 0, a.g.G - 1, a.g.G - 1,
                    [&](int32_t ti) {
                      pair_cell[o] = (int32_t)k;
                      pair_tri[o] = ti;
                      o++;
                    });
}

// ---------------------------------------------------------------------------
// Streaming Sutherland-Hodgman stage: clips by the closed half-plane left of
// a -> b.  Each vertex carries the label of its outgoing edge.
// ---------------------------------------------------------------------------
struct SV {
  double x, y;
  int lab;
};
struct Stage {
  double ax, ay, ex, ey, elen;
  bool started;
  SV first, prev;
  double fs, ps;
  __device__ void init(double x0, double y0, double x1, double y1) {
    ax = x0;
    ay = y0;
    ex = x1 - x0;
    ey = y1 - y0;
    elen = sqrt(ex * ex + ey * ey);
    started = false;
  }
  // Side of v w.r.t. the edge line; values within rounding of zero count as
  // on the line (inside: the triangle is closed), so a cell edge lying on a
  // mesh edge stays with the triangle on the cell's side.
  __device__ double side(const SV &v) const {
    const double dx = v.x - ax, dy = v.y - ay;
    const double s = ex * dy - ey * dx;
    return fabs(s) <= kSnap * elen * (fabs(v.x) + fabs(v.y) + fabs(dx) + fabs(dy)) ? 0.0 : s;
  }
  __device__ static SV cross_pt(const SV &p, double sp, const SV &c, double sc, int lab) {
    const double tt = sp / (sp - sc);
    SV r;
    r.x = p.x + tt * (c.x - p.x);
    r.y = p.y + tt * (c.y - p.y);
    r.lab = lab;
    return r;
  }
  // edge p -> c: emits up to two vertices
  __device__ int edge(const SV &p, double sp, const SV &c, double sc, SV &o0, SV &o1) const {
    if (sp >= 0.0) {
      if (sc >= 0.0) {
        o0 = c;
        return 1;
      }
      o0 = cross_pt(p, sp, c, sc, -1);
      return 1;
    }
    if (sc >= 0.0) {
      o0 = cross_pt(p, sp, c, sc, p.lab);
      o1 = c;
      return 2;
    }
    return 0;
  }
  __device__ int push(const SV &c, SV &o0, SV &o1) {
    const double s = side(c);
    if (!started) {
      started = true;
      first = c;
      fs = s;
      prev = c;
      ps = s;
      return 0;
    }
    const int m = edge(prev, ps, c, s, o0, o1);
    prev = c;
    ps = s;
    return m;
  }
  __device__ int flush(SV &o0, SV &o1) const {
    if (!started) return 0;
    return edge(prev, ps, first, fs, o0, o1);
  }
};

template <int MODE>
struct Integrator {
  // fragment vertices arrive in order; fan from the first one
  int cnt;
  SV first, prev;
  double area, mass, intg, mx, my;
  double r0, ga, gb;  // rho(Dirac + X) = r0 + ga X + gb Y
  // extent of the fragment, for the degeneracy test
  double xlo, xhi, ylo, yhi, cmax;
  // A fragment counts only if its area is positive: the reference clips the
  // closed triangle by the open cell (ties excluded), so a cell touching a
  // triangle only along an edge or at a point yields nothing.  Exactly
  // degenerate fragments come out with areas at rounding level, hence the
  // scale-aware threshold.  Edge integrals are buffered until the area is
  // known (named registers rather than arrays, so nothing spills to scratch).
  int nb;
  int bl0, bl1, bl2, bl3;
  double bv0, bv1, bv2, bv3;
  __device__ bool positive() const {
    return area > 1e-13 * ((xhi - xlo) + (yhi - ylo)) * cmax;
  }
  double *eacc;
  int32_t *ehit;
  int64_t N, k;
  __device__ void commit(int lab, double r) {
    atomicAdd(&eacc[lab * N + k], r);
    ehit[lab * N + k] = 1;
  }
  __device__ void edge(const SV &a, const SV &b) {
    if (MODE != 0 || a.lab < 0) return;
    const double ex = b.x - a.x, ey = b.y - a.y;
    if (ex == 0.0 && ey == 0.0) return;  // zero-length piece from a snapped vertex
    const double r =
        sqrt(ex * ex + ey * ey) * (r0 + ga * 0.5 * (a.x + b.x) + gb * 0.5 * (a.y + b.y));
    if (nb == 0) {
      bl0 = a.lab;
      bv0 = r;
    } else if (nb == 1) {
      bl1 = a.lab;
      bv1 = r;
    } else if (nb == 2) {
      bl2 = a.lab;
      bv2 = r;
    } else if (nb == 3) {
      bl3 = a.lab;
      bv3 = r;
    } else {
      // a degenerate fragment has at most one cell edge of positive length
      // (plus zero-length edges at tie vertices): five or more means a
      // genuine fragment
      commit(a.lab, r);
    }
    nb++;
  }
  __device__ void fan(const SV &p0, const SV &p1, const SV &p2) {
    const double x0 = p0.x, y0 = p0.y, x1 = p1.x, y1 = p1.y, x2 = p2.x, y2 = p2.y;
    const double ar = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
    const double gx = (x0 + x1 + x2) / 3.0, gy = (y0 + y1 + y2) / 3.0;
    area += ar;
    mass += ar * (r0 + ga * gx + gb * gy);
    AC6 q;
    ac6_points(x0, y0, x1, y1, x2, y2, q);
    double a1 = 0.0, a9 = 0.0, b1 = 0.0, b9 = 0.0;
#pragma unroll
    for (int p = 0; p < 6; p++) {
      const double rp = r0 + ga * q.px[p] + gb * q.py[p];
      if (MODE == 0) {
        const double f = rp * (q.px[p] * q.px[p] + q.py[p] * q.py[p]);
        if (p < 3) a1 += f; else a9 += f;
      } else {
        if (p < 3) {
          a1 += rp * q.px[p];
          b1 += rp * q.py[p];
        } else {
          a9 += rp * q.px[p];
          b9 += rp * q.py[p];
        }
      }
    }
    if (MODE == 0) {
      intg += ar * (a1 * (1.0 / 30.0) + a9 * (9.0 / 30.0));
    } else {
      mx += ar * (a1 * (1.0 / 30.0) + a9 * (9.0 / 30.0));
      my += ar * (b1 * (1.0 / 30.0) + b9 * (9.0 / 30.0));
    }
  }
  __device__ void push(const SV &v) {
    xlo = fmin(xlo, v.x);
    xhi = fmax(xhi, v.x);
    ylo = fmin(ylo, v.y);
    yhi = fmax(yhi, v.y);
    cmax = fmax(cmax, fmax(fabs(v.x), fabs(v.y)));
    if (cnt == 0) {
      first = v;
    } else {
      edge(prev, v);
      if (cnt >= 2) fan(first, prev, v);
    }
    prev = v;
    cnt++;
  }
  __device__ void finish() {
    if (cnt >= 2) edge(prev, first);
    if (positive()) {
      if (nb > 0) commit(bl0, bv0);
      if (nb > 1) commit(bl1, bv1);
      if (nb > 2) commit(bl2, bv2);
      if (nb > 3) commit(bl3, bv3);
    }
  }
};

template <int MODE>
__global__ void __launch_bounds__(256) k_frag(CellArgs a, int64_t npairs,
                                              const int32_t *__restrict__ pair_cell,
                                              const int32_t *__restrict__ pair_tri,
                                              unsigned long long *counters) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int64_t k = pair_cell[p];
  const TriData t = a.tri[pair_tri[p]];
  const int64_t N = a.n;
  const int n = a.pn[k];
  const double yx = a.pos[2 * k], yy = a.pos[2 * k + 1];
  Stage s1, s2, s3;
  s1.init(t.x[0] - yx, t.y[0] - yy, t.x[1] - yx, t.y[1] - yy);
  s2.init(t.x[1] - yx, t.y[1] - yy, t.x[2] - yx, t.y[2] - yy);
  s3.init(t.x[2] - yx, t.y[2] - yy, t.x[0] - yx, t.y[0] - yy);
  Integrator<MODE> I;
  I.cnt = 0;
  I.area = I.mass = I.intg = I.mx = I.my = 0.0;
  I.r0 = t.a * yx + t.b * yy + t.c;
  I.ga = t.a;
  I.gb = t.b;
  I.nb = 0;
  I.xlo = I.ylo = kInf;
  I.xhi = I.yhi = -kInf;
  I.cmax = 0.0;
  I.eacc = a.eacc;
  I.ehit = a.ehit;
  I.N = N;
  I.k = k;
  SV o0, o1, q0, q1, r0, r1;
  auto to3 = [&](const SV &v) {
    const int m3 = s3.push(v, r0, r1);
    if (m3 > 0) I.push(r0);
    if (m3 > 1) I.push(r1);
  };
  auto to2 = [&](const SV &v) {
    const int m2 = s2.push(v, q0, q1);
    if (m2 > 0) to3(q0);
    if (m2 > 1) to3(q1);
  };
  for (int v = 0; v < n; v++) {
    SV c;
    c.x = a.px[v * N + k];
    c.y = a.py[v * N + k];
    c.lab = (a.plab[v * N + k] >= 0) ? v : -1;
    const int m1 = s1.push(c, o0, o1);
    if (m1 > 0) to2(o0);
    if (m1 > 1) to2(o1);
  }
  {
    const int m1 = s1.flush(o0, o1);
    if (m1 > 0) to2(o0);
    if (m1 > 1) to2(o1);
    const int m2 = s2.flush(q0, q1);
    if (m2 > 0) to3(q0);
    if (m2 > 1) to3(q1);
    const int m3 = s3.flush(r0, r1);
    if (m3 > 0) I.push(r0);
    if (m3 > 1) I.push(r1);
    I.finish();
  }
  if (I.cnt >= 3 && I.positive()) {
    if (MODE == 0) {
      atomicAdd(&a.mass[k], I.mass);
      atomicAdd(&a.intg[k], I.intg);
    } else {
      atomicAdd(&a.mass[k], I.mass);
      atomicAdd(&a.mom[2 * k], I.mx);
      atomicAdd(&a.mom[2 * k + 1], I.my);
    }
    atomicAdd(&counters[0], 1ull);
  }
}

// Hessian row of every fast-path cell (include/MA/kantorovich.hpp:109-122):
// H_kj = -int_{e_kj} rho ds / (2 |y_k - y_j|), H_kk = -sum_j H_kj.
__global__ void k_rows(CellArgs a, double *fval_part, int32_t *row_cnt, int32_t *row_col,
                       double *row_val, int cap_row, unsigned long long *counters) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int n = (k < a.n) ? a.pn[k] : -1;  // n < 0: the general kernel writes this row
  unsigned long long my_max = 0, my_empty = 0;
  if (n >= 0) {
  const int64_t N = a.n;
  const double yx = a.pos[2 * k], yy = a.pos[2 * k + 1];
  int32_t cols[kMaxFast + 1];
  double vals[kMaxFast + 1];
  int nr = 0;
  double diag = 0.0;
  for (int v = 0; v < n; v++) {
    const int32_t j = a.plab[v * N + k];
    if (j < 0 || !a.ehit[v * N + k]) continue;
    const double dx = yx - a.pos[2 * j], dy = yy - a.pos[2 * j + 1];
    const double val = a.eacc[v * N + k] / (2.0 * sqrt(dx * dx + dy * dy));
    diag += val;
    int p = 0;
    while (p < nr && cols[p] < j) p++;
    if (p < nr && cols[p] == j) {
      vals[p] -= val;
      continue;
    }
    for (int r = nr; r > p; r--) {
      cols[r] = cols[r - 1];
      vals[r] = vals[r - 1];
    }
    cols[p] = j;
    vals[p] = -val;
    nr++;
  }
  int cnt = 0;
  if (nr > 0) {
    int32_t *oc = row_col + k * (int64_t)cap_row;
    double *ov = row_val + k * (int64_t)cap_row;
    bool placed = false;
    for (int r = 0; r < nr; r++) {
      if (!placed && cols[r] > k) {
        oc[cnt] = (int32_t)k;
        ov[cnt++] = diag;
        placed = true;
      }
      oc[cnt] = cols[r];
      ov[cnt++] = vals[r];
    }
    if (!placed) {
      oc[cnt] = (int32_t)k;
      ov[cnt++] = diag;
    }
  }
  row_cnt[k] = cnt;
  const double m = a.mass[k];
  fval_part[k] = a.w[k] * m - a.intg[k];
  my_max = (unsigned long long)n;
  my_empty = (m <= 0.0) ? 1ull : 0ull;
  }
  __shared__ unsigned long long s_max, s_empty;
  if (threadIdx.x == 0) {
    s_max = 0;
    s_empty = 0;
  }
  __syncthreads();
  atomicMax(&s_max, my_max);
  atomicAdd(&s_empty, my_empty);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicMax(&counters[2], s_max);
    atomicAdd(&counters[3], s_empty);
  }
}

// first moments: local (relative to the Dirac) -> absolute
__global__ void k_mom_finalize(CellArgs a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= a.n || a.pn[k] < 0) return;
  const double m = a.mass[k];
  a.mom[2 * k] += a.pos[2 * k] * m;
  a.mom[2 * k + 1] += a.pos[2 * k + 1] * m;
}

// Lists the cells whose vertex count holds the marker `value`.
__global__ void k_collect(int64_t n, const int32_t *pn, int32_t value, int32_t *list,
                          int32_t *count) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (pn[k] == value) list[atomicAdd(count, 1)] = (int32_t)k;
}

}  // namespace ma

namespace ma {

ma_status run_bin_candidates(ma_engine *e, const int32_t *pn);

ma_status run_cells(ma_engine *e, int mode, int32_t *npairs_out, int32_t *ngeneral_out) {
  EvalBuffers &B = e->eb;
  const int64_t n = e->n;
  MA_HIP_TRY(B.px.ensure((size_t)kMaxFast * n));
  MA_HIP_TRY(B.py.ensure((size_t)kMaxFast * n));
  MA_HIP_TRY(B.eacc.ensure((size_t)kMaxFast * n));
  MA_HIP_TRY(B.plab.ensure((size_t)kMaxFast * n));
  MA_HIP_TRY(B.ehit.ensure((size_t)kMaxFast * n));
  MA_HIP_TRY(B.pn.ensure((size_t)n));
  MA_HIP_TRY(B.bbox.ensure((size_t)(4 * n)));
  MA_HIP_TRY(B.tcount.ensure((size_t)n + 1));
  MA_HIP_TRY(B.poff.ensure((size_t)n + 1));
  MA_HIP_TRY(B.glist.ensure((size_t)n + 1));
  CellArgs a;
  a.g = e->grid;
  a.n = n;
  a.pos = e->pos.p;
  a.w = B.w_s.p;
  a.perm = e->perm.p;
  a.rect = B.rect.p;
  a.cand_cnt = B.cand_cnt.p;
  a.cand = B.cand.p;
  a.cap = e->cap_cand;
  a.tbin_start = e->tbin_start.p;
  a.tbin_items = e->tbin_items.p;
  a.tri = e->tri.p;
  a.px = B.px.p;
  a.py = B.py.p;
  a.plab = B.plab.p;
  a.pn = B.pn.p;
  a.bbox = B.bbox.p;
  a.tcount = B.tcount.p;
  a.mass = B.mass_s.p;
  a.intg = B.intg_s.p;
  a.mom = B.mom_s.p;
  a.eacc = B.eacc.p;
  a.ehit = B.ehit.p;
  a.flags = B.flags.p;
  a.force_general = e->force_general;
  a.bin_start = e->bin_start.p;
  for (int i = 0; i < 4; i++) a.dom[i] = e->dom[i];
  const int32_t zero = 0;
  hipLaunchKernelGGL(k_cell_local, dim3(nblocks(n, kCellBlock)), dim3(kCellBlock), 0, e->stream,
                     a, B.pyr.p + e->pyr_off[e->grid.L], kLocalRings);
  {
    int32_t fl0[4];
    MA_HIP_TRY(hipMemcpyAsync(fl0, B.flags.p, sizeof fl0, hipMemcpyDeviceToHost, e->stream));
    MA_HIP_TRY(hipStreamSynchronize(e->stream));
    e->info.nbinpath = 0;
    if (fl0[2] || fl0[3]) {
      // cells the site-centric search could not certify: bin-candidate path
      MA_HIP_TRY(hipEventRecord(e->ev[8], e->stream));
      ma_status s = run_bin_candidates(e, B.pn.p);
      if (s != MA_OK) return s;
      a.cand = B.cand.p;
      a.cap = e->cap_cand;
      MA_HIP_TRY(hipMemcpyAsync(B.glist.p + n, &zero, sizeof(int32_t), hipMemcpyHostToDevice,
                                e->stream));
      hipLaunchKernelGGL(k_collect, dim3(nblocks(n)), dim3(kBlock), 0, e->stream, n, B.pn.p, -2,
                         B.glist.p, B.glist.p + n);
      int32_t nb = 0;
      MA_HIP_TRY(hipMemcpyAsync(&nb, B.glist.p + n, sizeof(int32_t), hipMemcpyDeviceToHost,
                                e->stream));
      MA_HIP_TRY(hipStreamSynchronize(e->stream));
      e->info.nbinpath = nb;
      if (nb > 0)
        hipLaunchKernelGGL(k_cell_bins, dim3(nblocks(nb, kCellBlock)), dim3(kCellBlock), 0,
                           e->stream, a, B.glist.p, B.glist.p + n);
      MA_HIP_TRY(hipEventRecord(e->ev[9], e->stream));
    }
  }
  MA_HIP_TRY(hipEventRecord(e->ev[3], e->stream));
  MA_HIP_TRY(hipMemcpyAsync(B.tcount.p + n, &zero, sizeof(int32_t), hipMemcpyHostToDevice,
                            e->stream));
  MA_HIP_TRY(exclusive_scan(e, B.tcount.p, B.poff.p, n + 1));
  int32_t np = 0, fl[4];
  MA_HIP_TRY(hipMemcpyAsync(&np, B.poff.p + n, sizeof(int32_t), hipMemcpyDeviceToHost,
                            e->stream));
  MA_HIP_TRY(hipMemcpyAsync(fl, B.flags.p, sizeof fl, hipMemcpyDeviceToHost, e->stream));
  MA_HIP_TRY(hipStreamSynchronize(e->stream));
  MA_HIP_TRY(B.pair_cell.ensure((size_t)np + 1));
  MA_HIP_TRY(B.pair_tri.ensure((size_t)np + 1));
  hipLaunchKernelGGL(k_pairs, dim3(nblocks(n)), dim3(kBlock), 0, e->stream, a, B.poff.p,
                     B.pair_cell.p, B.pair_tri.p);
  MA_HIP_TRY(hipEventRecord(e->ev[4], e->stream));
  if (np > 0) {
    if (mode == 0)
      hipLaunchKernelGGL(k_frag<0>, dim3(nblocks(np, 256)), dim3(256), 0, e->stream, a,
                         (int64_t)np, B.pair_cell.p, B.pair_tri.p, B.counters.p);
    else
      hipLaunchKernelGGL(k_frag<1>, dim3(nblocks(np, 256)), dim3(256), 0, e->stream, a,
                         (int64_t)np, B.pair_cell.p, B.pair_tri.p, B.counters.p);
  }
  MA_HIP_TRY(hipEventRecord(e->ev[5], e->stream));
  if (mode == 0)
    hipLaunchKernelGGL(k_rows, dim3(nblocks(n)), dim3(kBlock), 0, e->stream, a, B.fval_part.p,
                       B.row_cnt.p, B.row_col.p, B.row_val.p, e->cap_row, B.counters.p);
  else
    hipLaunchKernelGGL(k_mom_finalize, dim3(nblocks(n)), dim3(kBlock), 0, e->stream, a);
  int32_t ng = 0;
  if (fl[3]) {
    MA_HIP_TRY(hipMemcpyAsync(B.glist.p + n, &zero, sizeof(int32_t), hipMemcpyHostToDevice,
                              e->stream));
    hipLaunchKernelGGL(k_collect, dim3(nblocks(n)), dim3(kBlock), 0, e->stream, n, B.pn.p, -1,
                       B.glist.p, B.glist.p + n);
    MA_HIP_TRY(hipMemcpyAsync(&ng, B.glist.p + n, sizeof(int32_t), hipMemcpyDeviceToHost,
                              e->stream));
    MA_HIP_TRY(hipStreamSynchronize(e->stream));
  }
  MA_HIP_TRY(hipGetLastError());
  *npairs_out = np;
  *ngeneral_out = ng;
  return MA_OK;
}

}  // namespace ma
