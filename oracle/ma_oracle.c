/*
 * CPU oracle (FP64) for the power-diagram x triangulation hot path.
 * TEST INFRASTRUCTURE ONLY -- see ma_oracle.h for scope and citations.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off; no fast-math, the
 * exact predicates rely on IEEE round-to-nearest-even error-free transforms).
 */
#include "ma_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ====================================================================== */
/* Error-free transforms and floating-point expansions.                    */
/* Published method: J. R. Shewchuk, "Adaptive Precision Floating-Point    */
/* Arithmetic and Fast Robust Geometric Predicates", DCG 18 (1997).        */
/* ====================================================================== */

static inline void two_sum(double a, double b, double *x, double *y) {
  double s = a + b, bv = s - a, av = s - bv;
  *y = (a - av) + (b - bv);
  *x = s;
}
static inline void fast_two_sum(double a, double b, double *x, double *y) {
  double s = a + b;
  *y = b - (s - a);
  *x = s;
}
static inline void two_diff(double a, double b, double *x, double *y) {
  double s = a - b, bv = a - s, av = s + bv;
  *y = (a - av) + (bv - b);
  *x = s;
}
static inline void two_prod(double a, double b, double *x, double *y) {
  double p = a * b;
  *y = fma(a, b, -p);
  *x = p;
}

#define XM 640
typedef struct {
  int n;
  double t[XM];
} xp; /* nonoverlapping expansion, increasing magnitude, zero-eliminated */

static void xp_set(xp *r, double a) {
  r->n = 1;
  r->t[0] = a;
}
static void xp_diff(xp *r, double a, double b) {
  double x, y;
  two_diff(a, b, &x, &y);
  if (y != 0.0) {
    r->n = 2;
    r->t[0] = y;
    r->t[1] = x;
  } else {
    r->n = 1;
    r->t[0] = x;
  }
}
/* h = e + b (grow_expansion_zeroelim); h must not alias e */
static int grow_ze(int elen, const double *e, double b, double *h) {
  double Q = b, Qn, hh;
  int k = 0;
  for (int i = 0; i < elen; i++) {
    two_sum(Q, e[i], &Qn, &hh);
    Q = Qn;
    if (hh != 0.0) h[k++] = hh;
  }
  if (Q != 0.0 || k == 0) h[k++] = Q;
  return k;
}
/* h = e * b (scale_expansion_zeroelim) */
static int scale_ze(int elen, const double *e, double b, double *h) {
  double Q, sum, hh, p1, p0;
  int k = 0;
  two_prod(e[0], b, &Q, &hh);
  if (hh != 0.0) h[k++] = hh;
  for (int i = 1; i < elen; i++) {
    two_prod(e[i], b, &p1, &p0);
    two_sum(Q, p0, &sum, &hh);
    if (hh != 0.0) h[k++] = hh;
    fast_two_sum(p1, sum, &Q, &hh);
    if (hh != 0.0) h[k++] = hh;
  }
  if (Q != 0.0 || k == 0) h[k++] = Q;
  return k;
}
/* in-place compress (Shewchuk's compress) */
static void xp_compress(xp *e) {
  double h[XM];
  int elen = e->n;
  int bottom = elen - 1;
  double Q = e->t[bottom], q, Qn;
  for (int i = elen - 2; i >= 0; i--) {
    fast_two_sum(Q, e->t[i], &Qn, &q);
    if (q != 0.0) {
      h[bottom--] = Qn;
      Q = q;
    } else {
      Q = Qn;
    }
  }
  int top = 0;
  for (int i = bottom + 1; i < elen; i++) {
    fast_two_sum(h[i], Q, &Qn, &q);
    if (q != 0.0) h[top++] = q;
    Q = Qn;
  }
  h[top] = Q;
  e->n = top + 1;
  memcpy(e->t, h, sizeof(double) * (size_t)e->n);
}
static void xp_add(xp *r, const xp *a, const xp *b) {
  double t1[XM], t2[XM];
  int n = a->n;
  memcpy(t1, a->t, sizeof(double) * (size_t)n);
  double *cur = t1, *nxt = t2;
  for (int i = 0; i < b->n; i++) {
    n = grow_ze(n, cur, b->t[i], nxt);
    double *s = cur;
    cur = nxt;
    nxt = s;
  }
  r->n = n;
  memcpy(r->t, cur, sizeof(double) * (size_t)n);
  xp_compress(r);
}
static void xp_neg(xp *r) {
  for (int i = 0; i < r->n; i++) r->t[i] = -r->t[i];
}
static void xp_sub(xp *r, const xp *a, const xp *b) {
  xp nb = *b;
  xp_neg(&nb);
  xp_add(r, a, &nb);
}
static void xp_mul(xp *r, const xp *a, const xp *b) {
  xp acc, s;
  xp_set(&acc, 0.0);
  for (int k = 0; k < b->n; k++) {
    s.n = scale_ze(a->n, a->t, b->t[k], s.t);
    xp tmp;
    xp_add(&tmp, &acc, &s);
    acc = tmp;
  }
  *r = acc;
}
static int xp_sign(const xp *a) {
  double top = a->t[a->n - 1];
  return (top > 0) - (top < 0);
}

/* Running-error-bound arithmetic for the floating-point filter stage. */
typedef struct {
  double v, e;
} ev;
#define EPS_U 1.1102230246251565e-16
static inline ev ev_c(double x) {
  ev r = {x, 0.0};
  return r;
}
static inline ev ev_add(ev a, ev b) {
  ev r;
  r.v = a.v + b.v;
  r.e = a.e + b.e + EPS_U * fabs(r.v);
  return r;
}
static inline ev ev_sub(ev a, ev b) {
  ev r;
  r.v = a.v - b.v;
  r.e = a.e + b.e + EPS_U * fabs(r.v);
  return r;
}
static inline ev ev_mul(ev a, ev b) {
  ev r;
  r.v = a.v * b.v;
  r.e = fabs(a.v) * b.e + fabs(b.v) * a.e + a.e * b.e + EPS_U * fabs(r.v);
  return r;
}
/* -1/0/+1 if certain, 2 if uncertain */
static inline int ev_sign(ev a) {
  double bound = a.e * (1.0 + 1e-10) + 1e-300;
  if (a.v > bound) return 1;
  if (a.v < -bound) return -1;
  return 2;
}

static int64_t g_npred = 0, g_nexact = 0;

/* ---- Side1: sign(pi_v(p) - pi_w(p)) at an input point p
 *      (predicates.hpp:73-99, Compare_power_distance_2) ---- */
int orc_side1_exact(const double *v, const double *w, double px, double py) {
  xp a, b, s1, s2, t, acc, acc2, cw;
  xp_diff(&a, px, v[0]);
  xp_diff(&b, py, v[1]);
  xp_mul(&s1, &a, &a);
  xp_mul(&s2, &b, &b);
  xp_add(&acc, &s1, &s2); /* |p-v|^2 */
  xp_diff(&a, px, w[0]);
  xp_diff(&b, py, w[1]);
  xp_mul(&s1, &a, &a);
  xp_mul(&s2, &b, &b);
  xp_add(&acc2, &s1, &s2); /* |p-w|^2 */
  xp_sub(&t, &acc, &acc2);
  xp_diff(&cw, w[2], v[2]); /* - v_w + w_w */
  xp_add(&acc, &t, &cw);
  return xp_sign(&acc);
}
int orc_side1(const double *v, const double *w, double px, double py) {
  g_npred++;
  ev ax = ev_sub(ev_c(px), ev_c(v[0])), ay = ev_sub(ev_c(py), ev_c(v[1]));
  ev bx = ev_sub(ev_c(px), ev_c(w[0])), by = ev_sub(ev_c(py), ev_c(w[1]));
  ev d = ev_add(ev_mul(ax, ax), ev_mul(ay, ay));
  d = ev_sub(d, ev_add(ev_mul(bx, bx), ev_mul(by, by)));
  d = ev_add(d, ev_sub(ev_c(w[2]), ev_c(v[2])));
  int s = ev_sign(d);
  if (s != 2) return s;
  g_nexact++;
  return orc_side1_exact(v, w, px, py);
}

/* ---- Side2: sign of pi_v - pi_w at the power vertex of (v,u2,u3)
 *      (predicates.hpp:101-115).  With U=u-v, k_u=|U|^2+v_w-u_w:
 *      pi_v-pi_w at c = num/det, num = k2 x(W,U3) + k3 x(U2,W) + kw x(U3,U2),
 *      det = x(U2,U3). ---- */
static void xp_k(xp *k, const xp *Ux, const xp *Uy, double vw, double uw) {
  xp s1, s2, t, c;
  xp_mul(&s1, Ux, Ux);
  xp_mul(&s2, Uy, Uy);
  xp_add(&t, &s1, &s2);
  xp_diff(&c, vw, uw);
  xp_add(k, &t, &c);
}
static void xp_cross(xp *r, const xp *ax, const xp *ay, const xp *bx, const xp *by) {
  xp s1, s2;
  xp_mul(&s1, ax, by);
  xp_mul(&s2, ay, bx);
  xp_sub(r, &s1, &s2);
}
int orc_side2_exact(const double *v, const double *u2, const double *u3, const double *w) {
  xp U2x, U2y, U3x, U3y, Wx, Wy, k2, k3, kw, c1, c2, c3, det, t1, t2, t3, s, num;
  xp_diff(&U2x, u2[0], v[0]);
  xp_diff(&U2y, u2[1], v[1]);
  xp_diff(&U3x, u3[0], v[0]);
  xp_diff(&U3y, u3[1], v[1]);
  xp_diff(&Wx, w[0], v[0]);
  xp_diff(&Wy, w[1], v[1]);
  xp_k(&k2, &U2x, &U2y, v[2], u2[2]);
  xp_k(&k3, &U3x, &U3y, v[2], u3[2]);
  xp_k(&kw, &Wx, &Wy, v[2], w[2]);
  xp_cross(&c1, &Wx, &Wy, &U3x, &U3y);
  xp_cross(&c2, &U2x, &U2y, &Wx, &Wy);
  xp_cross(&c3, &U3x, &U3y, &U2x, &U2y);
  xp_cross(&det, &U2x, &U2y, &U3x, &U3y);
  xp_mul(&t1, &k2, &c1);
  xp_mul(&t2, &k3, &c2);
  xp_mul(&t3, &kw, &c3);
  xp_add(&s, &t1, &t2);
  xp_add(&num, &s, &t3);
  int sd = xp_sign(&det);
  if (sd == 0) return 0; /* degenerate: no vertex; treated as a tie */
  return xp_sign(&num) * sd;
}
int orc_side2(const double *v, const double *u2, const double *u3, const double *w) {
  g_npred++;
  ev U2x = ev_sub(ev_c(u2[0]), ev_c(v[0])), U2y = ev_sub(ev_c(u2[1]), ev_c(v[1]));
  ev U3x = ev_sub(ev_c(u3[0]), ev_c(v[0])), U3y = ev_sub(ev_c(u3[1]), ev_c(v[1]));
  ev Wx = ev_sub(ev_c(w[0]), ev_c(v[0])), Wy = ev_sub(ev_c(w[1]), ev_c(v[1]));
  ev k2 = ev_add(ev_add(ev_mul(U2x, U2x), ev_mul(U2y, U2y)), ev_sub(ev_c(v[2]), ev_c(u2[2])));
  ev k3 = ev_add(ev_add(ev_mul(U3x, U3x), ev_mul(U3y, U3y)), ev_sub(ev_c(v[2]), ev_c(u3[2])));
  ev kw = ev_add(ev_add(ev_mul(Wx, Wx), ev_mul(Wy, Wy)), ev_sub(ev_c(v[2]), ev_c(w[2])));
  ev c1 = ev_sub(ev_mul(Wx, U3y), ev_mul(Wy, U3x));
  ev c2 = ev_sub(ev_mul(U2x, Wy), ev_mul(U2y, Wx));
  ev c3 = ev_sub(ev_mul(U3x, U2y), ev_mul(U3y, U2x));
  ev det = ev_sub(ev_mul(U2x, U3y), ev_mul(U2y, U3x));
  ev num = ev_add(ev_add(ev_mul(k2, c1), ev_mul(k3, c2)), ev_mul(kw, c3));
  int sn = ev_sign(num), sd = ev_sign(det);
  if (sn != 2 && sd != 2 && sd != 0) return sn * sd;
  g_nexact++;
  return orc_side2_exact(v, u2, u3, w);
}

/* ---- Side3: sign of pi_v - pi_w at line(p,q) cap radical_axis(v,u)
 *      (predicates.hpp:117-135).  With P=p-v, E=q-p:
 *      num = (2P.W - k_w)(E.U) + (k_u - 2P.U)(E.W), den = E.U. ---- */
int orc_side3_exact(const double *v, const double *u, double px, double py, double qx,
                    double qy, const double *w) {
  xp Px, Py, Ex, Ey, Ux, Uy, Wx, Wy, ku, kw, EU, EW, PU, PW, t1, t2, a, b, num, two;
  xp_diff(&Px, px, v[0]);
  xp_diff(&Py, py, v[1]);
  xp_diff(&Ex, qx, px);
  xp_diff(&Ey, qy, py);
  xp_diff(&Ux, u[0], v[0]);
  xp_diff(&Uy, u[1], v[1]);
  xp_diff(&Wx, w[0], v[0]);
  xp_diff(&Wy, w[1], v[1]);
  xp_k(&ku, &Ux, &Uy, v[2], u[2]);
  xp_k(&kw, &Wx, &Wy, v[2], w[2]);
  xp_mul(&t1, &Ex, &Ux);
  xp_mul(&t2, &Ey, &Uy);
  xp_add(&EU, &t1, &t2);
  xp_mul(&t1, &Ex, &Wx);
  xp_mul(&t2, &Ey, &Wy);
  xp_add(&EW, &t1, &t2);
  xp_set(&two, 2.0);
  xp_mul(&t1, &Px, &Ux);
  xp_mul(&t2, &Py, &Uy);
  xp_add(&a, &t1, &t2);
  xp_mul(&PU, &a, &two);
  xp_mul(&t1, &Px, &Wx);
  xp_mul(&t2, &Py, &Wy);
  xp_add(&a, &t1, &t2);
  xp_mul(&PW, &a, &two);
  xp_sub(&a, &PW, &kw);
  xp_mul(&t1, &a, &EU);
  xp_sub(&b, &ku, &PU);
  xp_mul(&t2, &b, &EW);
  xp_add(&num, &t1, &t2);
  int sd = xp_sign(&EU);
  if (sd == 0) return 0;
  return xp_sign(&num) * sd;
}
int orc_side3(const double *v, const double *u, double px, double py, double qx,
              double qy, const double *w) {
  g_npred++;
  ev Px = ev_sub(ev_c(px), ev_c(v[0])), Py = ev_sub(ev_c(py), ev_c(v[1]));
  ev Ex = ev_sub(ev_c(qx), ev_c(px)), Ey = ev_sub(ev_c(qy), ev_c(py));
  ev Ux = ev_sub(ev_c(u[0]), ev_c(v[0])), Uy = ev_sub(ev_c(u[1]), ev_c(v[1]));
  ev Wx = ev_sub(ev_c(w[0]), ev_c(v[0])), Wy = ev_sub(ev_c(w[1]), ev_c(v[1]));
  ev ku = ev_add(ev_add(ev_mul(Ux, Ux), ev_mul(Uy, Uy)), ev_sub(ev_c(v[2]), ev_c(u[2])));
  ev kw = ev_add(ev_add(ev_mul(Wx, Wx), ev_mul(Wy, Wy)), ev_sub(ev_c(v[2]), ev_c(w[2])));
  ev EU = ev_add(ev_mul(Ex, Ux), ev_mul(Ey, Uy));
  ev EW = ev_add(ev_mul(Ex, Wx), ev_mul(Ey, Wy));
  ev PU = ev_mul(ev_c(2.0), ev_add(ev_mul(Px, Ux), ev_mul(Py, Uy)));
  ev PW = ev_mul(ev_c(2.0), ev_add(ev_mul(Px, Wx), ev_mul(Py, Wy)));
  ev num = ev_add(ev_mul(ev_sub(PW, kw), EU), ev_mul(ev_sub(ku, PU), EW));
  int sn = ev_sign(num), sd = ev_sign(EU);
  if (sn != 2 && sd != 2 && sd != 0) return sn * sd;
  g_nexact++;
  return orc_side3_exact(v, u, px, py, qx, qy, w);
}

/* ====================================================================== */
/* Mesh, density (functions.hpp:23-80) and quadrature (quadrature.hpp).    */
/* ====================================================================== */

typedef struct {
  double x[3], y[3]; /* counter-clockwise, as CGAL faces are */
  double a, b, c;    /* Linear_function: rho(x,y) = a x + b y + c */
} otri;

/* barycentric (functions.hpp:23-41) */
static void barycentric(double px, double py, double ax, double ay, double bx, double by,
                        double cx, double cy, double *u, double *v, double *w) {
  double v0x = bx - ax, v0y = by - ay, v1x = cx - ax, v1y = cy - ay, v2x = px - ax,
         v2y = py - ay;
  double d00 = v0x * v0x + v0y * v0y;
  double d01 = v0x * v1x + v0y * v1y;
  double d11 = v1x * v1x + v1y * v1y;
  double d20 = v2x * v0x + v2y * v0y;
  double d21 = v2x * v1x + v2y * v1y;
  double denom = d00 * d11 - d01 * d01;
  *v = (d11 * d20 - d01 * d21) / denom;
  *w = (d00 * d21 - d01 * d20) / denom;
  *u = 1.0 - *v - *w;
}
/* extrapolate (functions.hpp:43-53) */
static double extrapolate(double px, double py, const otri *t, const double *f) {
  double u, v, w;
  barycentric(px, py, t->x[0], t->y[0], t->x[1], t->y[1], t->x[2], t->y[2], &u, &v, &w);
  return u * f[0] + v * f[1] + w * f[2];
}
/* Linear_function constructor (functions.hpp:62-71) */
static void linear_function(otri *t, const double *f) {
  t->c = extrapolate(0.0, 0.0, t, f);
  t->a = extrapolate(1.0, 0.0, t, f) - t->c;
  t->b = extrapolate(0.0, 1.0, t, f) - t->c;
}
static inline double rho_at(const otri *t, double x, double y) {
  return t->a * x + t->b * y + t->c;
}
/* CGAL::area(a,b,c) */
static inline double tri_area(double ax, double ay, double bx, double by, double cx, double cy) {
  return ((bx - ax) * (cy - ay) - (cx - ax) * (by - ay)) / 2.0;
}
/* integrate_centroid (quadrature.hpp:69-77) */
static double integrate_centroid(double ax, double ay, double bx, double by, double cx,
                                 double cy, const otri *t) {
  double gx = (ax + bx + cx) / 3.0, gy = (ay + by + cy) / 3.0;
  return tri_area(ax, ay, bx, by, cx, cy) * rho_at(t, gx, gy);
}
/* Albrecht-Collatz 6-point rule (quadrature.hpp:23-42), integrand given as a
 * callback on (x,y) returning nout values. */
typedef void (*integrand_fn)(double x, double y, const void *ctx, double *out);
static void integrate_ac6(double ax, double ay, double bx, double by, double cx, double cy,
                          integrand_fn f, const void *ctx, int nout, double *res) {
  const double h = 1.0 / 2.0, s6 = 1.0 / 6.0, t3 = 2.0 / 3.0;
  const double w1 = 1.0 / 30.0, w9 = 9.0 / 30.0;
  double ux = bx - ax, uy = by - ay, vx = cx - ax, vy = cy - ay;
  double pts[6][2] = {{ax + h * ux + h * vx, ay + h * uy + h * vy},
                      {ax + h * ux, ay + h * uy},
                      {ax + h * vx, ay + h * vy},
                      {ax + s6 * ux + t3 * vx, ay + s6 * uy + t3 * vy},
                      {ax + s6 * vx + t3 * ux, ay + s6 * vy + t3 * uy},
                      {ax + s6 * ux + s6 * vx, ay + s6 * uy + s6 * vy}};
  double wts[6] = {w1, w1, w1, w9, w9, w9};
  double acc[8] = {0};
  for (int k = 0; k < 6; k++) {
    double o[8];
    f(pts[k][0], pts[k][1], ctx, o);
    for (int q = 0; q < nout; q++) acc[q] += wts[k] * o[q];
  }
  double A = tri_area(ax, ay, bx, by, cx, cy);
  for (int q = 0; q < nout; q++) res[q] = A * acc[q];
}

typedef struct {
  const otri *t;
  double yx, yy;
} kctx;
/* rho(x) |x - y_v|^2  (kantorovich.hpp:126-131) */
static void f_intg(double x, double y, const void *c, double *o) {
  const kctx *k = (const kctx *)c;
  double dx = x - k->yx, dy = y - k->yy;
  o[0] = rho_at(k->t, x, y) * (dx * dx + dy * dy);
}
/* (rho, rho x, rho y)  (lloyd.hpp:45-60) */
static void f_moment(double x, double y, const void *c, double *o) {
  const kctx *k = (const kctx *)c;
  double r = rho_at(k->t, x, y);
  o[0] = r;
  o[1] = r * x;
  o[2] = r * y;
}
/* (rho, rho x, rho y, rho x^2, rho y^2, rho x y)  (lloyd.hpp:101-112) */
static void f_moment2(double x, double y, const void *c, double *o) {
  const kctx *k = (const kctx *)c;
  double r = rho_at(k->t, x, y);
  o[0] = r;
  o[1] = r * x;
  o[2] = r * y;
  o[3] = r * x * x;
  o[4] = r * y * y;
  o[5] = r * x * y;
}

typedef struct {
  const double *c;
} pctx;
static void f_poly(double x, double y, const void *cv, double *o) {
  const double *c = ((const pctx *)cv)->c;
  o[0] = c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y +
         c[6] * x * x * x + c[7] * x * x * y + c[8] * x * y * y + c[9] * y * y * y;
}
double orc_ac6_poly(const double *t6, const double *c10) {
  pctx p = {c10};
  double r;
  integrate_ac6(t6[0], t6[1], t6[2], t6[3], t6[4], t6[5], f_poly, &p, 1, &r);
  return r;
}
double orc_centroid_lin(const double *t6, double a, double b, double c) {
  otri t;
  memset(&t, 0, sizeof t);
  t.a = a;
  t.b = b;
  t.c = c;
  return integrate_centroid(t6[0], t6[1], t6[2], t6[3], t6[4], t6[5], &t);
}

/* ====================================================================== */
/* Uniform bin grid.                                                       */
/* ====================================================================== */
typedef struct {
  double x0, y0, h;
  int gx, gy;
  int32_t *start, *items;
} ogrid;

static void grid_free(ogrid *g) {
  free(g->start);
  free(g->items);
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int bin_x(const ogrid *g, double x) {
  return clampi((int)floor((x - g->x0) / g->h), 0, g->gx - 1);
}
static inline int bin_y(const ogrid *g, double y) {
  return clampi((int)floor((y - g->y0) / g->h), 0, g->gy - 1);
}

/* ====================================================================== */
/* Power cells (FP64) -> neighbour supersets.  Replaces the CGAL regular   */
/* triangulation (kantorovich.hpp:64-72); certified by a security radius.  */
/* ====================================================================== */
#define MAXPV 512
typedef struct {
  int n;
  double x[MAXPV], y[MAXPV]; /* local coordinates, relative to the site */
  int lab[MAXPV];            /* label of edge k -> k+1 (neighbour or <0 box) */
} opoly;

/* keep L(X) = k - 2 X.Y > 0 ; returns 0 if ok, -1 if overflow */
static int cell_clip(opoly *P, double Yx, double Yy, double k, int label) {
  double s[MAXPV];
  int n = P->n, nin = 0;
  for (int i = 0; i < n; i++) {
    s[i] = k - 2.0 * (P->x[i] * Yx + P->y[i] * Yy);
    nin += s[i] > 0.0;
  }
  if (nin == n) return 0;
  if (nin == 0) {
    P->n = 0;
    return 0;
  }
  opoly Q;
  int m = 0;
  for (int i = 0; i < n; i++) {
    int i1 = (i + 1) % n;
    if (m + 2 > MAXPV) return -1;
    if (s[i] > 0.0) {
      Q.x[m] = P->x[i];
      Q.y[m] = P->y[i];
      Q.lab[m] = P->lab[i];
      m++;
      if (!(s[i1] > 0.0)) {
        double t = s[i] / (s[i] - s[i1]);
        Q.x[m] = P->x[i] + t * (P->x[i1] - P->x[i]);
        Q.y[m] = P->y[i] + t * (P->y[i1] - P->y[i]);
        Q.lab[m] = label;
        m++;
      }
    } else if (s[i1] > 0.0) {
      double t = s[i] / (s[i] - s[i1]);
      Q.x[m] = P->x[i] + t * (P->x[i1] - P->x[i]);
      Q.y[m] = P->y[i] + t * (P->y[i1] - P->y[i]);
      Q.lab[m] = P->lab[i];
      m++;
    }
  }
  Q.n = m;
  *P = Q;
  return 0;
}

typedef struct {
  int64_t n;
  const double *X, *w;
  double bx0, by0, bx1, by1; /* domain bounding box */
  double W;                  /* max weight */
  double delta;              /* near-touch distance */
  ogrid sg;                  /* site grid */
} cellctx;

/* Computes the cell of site i clipped to the domain box; appends neighbour
 * candidates (every j whose radical axis comes within delta of the FP64
 * cell) to nb.  Returns number of neighbours, or -1 on overflow. */
static int power_cell(const cellctx *C, int64_t i, opoly *P, int32_t *nb, int maxnb,
                      double *bbox) {
  const double xi = C->X[2 * i], yi = C->X[2 * i + 1], wi = C->w[i];
  P->n = 4;
  P->x[0] = C->bx0 - xi; P->y[0] = C->by0 - yi; P->lab[0] = -1;
  P->x[1] = C->bx1 - xi; P->y[1] = C->by0 - yi; P->lab[1] = -2;
  P->x[2] = C->bx1 - xi; P->y[2] = C->by1 - yi; P->lab[2] = -3;
  P->x[3] = C->bx0 - xi; P->y[3] = C->by1 - yi; P->lab[3] = -4;
  int nnb = 0;
  const ogrid *g = &C->sg;
  int bx = bin_x(g, xi), by = bin_y(g, yi);
  for (int r = 0;; r++) {
    double rho2 = 0.0;
    for (int k = 0; k < P->n; k++) {
      double d2 = P->x[k] * P->x[k] + P->y[k] * P->y[k];
      if (d2 > rho2) rho2 = d2;
    }
    double rho = sqrt(rho2);
    for (int yy = by - r; yy <= by + r; yy++) {
      if (yy < 0 || yy >= g->gy) continue;
      int ring_row = (yy == by - r || yy == by + r);
      for (int xx = bx - r; xx <= bx + r; xx++) {
        if (xx < 0 || xx >= g->gx) continue;
        if (!ring_row && xx != bx - r && xx != bx + r) continue;
        int b = yy * g->gx + xx;
        for (int32_t q = g->start[b]; q < g->start[b + 1]; q++) {
          int32_t j = g->items[q];
          if (j == i || P->n == 0) continue;
          double Yx = C->X[2 * j] - xi, Yy = C->X[2 * j + 1] - yi;
          double d2 = Yx * Yx + Yy * Yy;
          double d = sqrt(d2);
          if (d == 0.0) {
            /* coincident sites: the one with the larger weight hides the other */
            if (C->w[j] > wi || (C->w[j] == wi && j < i)) P->n = 0;
            continue;
          }
          double lb = d2 - 2.0 * rho * d + wi - C->w[j];
          if (lb / (2.0 * d) > C->delta) continue;
          double kk = d2 + wi - C->w[j];
          double mins = INFINITY;
          for (int k = 0; k < P->n; k++) {
            double sk = kk - 2.0 * (P->x[k] * Yx + P->y[k] * Yy);
            if (sk < mins) mins = sk;
          }
          if (mins / (2.0 * d) > C->delta) continue;
          if (nnb >= maxnb) return -1;
          nb[nnb++] = j;
          if (mins <= 0.0) {
            if (cell_clip(P, Yx, Yy, kk, j) < 0) return -1;
            rho2 = 0.0;
            for (int k = 0; k < P->n; k++) {
              double dd = P->x[k] * P->x[k] + P->y[k] * P->y[k];
              if (dd > rho2) rho2 = dd;
            }
            rho = sqrt(rho2);
          }
        }
      }
    }
    if (P->n == 0) break;
    /* certification: every unvisited site is at distance >= D from y_i */
    int x_lo = bx - r, x_hi = bx + r, y_lo = by - r, y_hi = by + r;
    double D = INFINITY;
    if (x_lo > 0) D = fmin(D, xi - (g->x0 + x_lo * g->h));
    if (x_hi < g->gx - 1) D = fmin(D, g->x0 + (x_hi + 1) * g->h - xi);
    if (y_lo > 0) D = fmin(D, yi - (g->y0 + y_lo * g->h));
    if (y_hi < g->gy - 1) D = fmin(D, g->y0 + (y_hi + 1) * g->h - yi);
    if (D == INFINITY) break;
    if (D > rho && (D * D - 2.0 * rho * D - (C->W - wi)) / (2.0 * D) > 2.0 * C->delta) break;
  }
  if (P->n > 0) {
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int k = 0; k < P->n; k++) {
      x0 = fmin(x0, P->x[k]); x1 = fmax(x1, P->x[k]);
      y0 = fmin(y0, P->y[k]); y1 = fmax(y1, P->y[k]);
    }
    double pad = 4.0 * C->delta;
    bbox[0] = x0 + xi - pad; bbox[1] = y0 + yi - pad;
    bbox[2] = x1 + xi + pad; bbox[3] = y1 + yi + pad;
  }
  return nnb;
}

/* ====================================================================== */
/* Tri_intersector (voronoi_triangulation_intersection.hpp:42-199).        */
/* ====================================================================== */
enum { EDGE_T = 0, EDGE_DT = 1 };
typedef struct {
  int type;
  int a, b; /* EDGE_T: triangle-local vertex indices (a,b); EDGE_DT: b = neighbour */
} pedge;
#define MAXPE 512

typedef struct {
  const otri *t;
  const double *X, *w;
  int64_t v;
} tictx;

static inline void wp(const tictx *c, int64_t i, double *o) {
  o[0] = c->X[2 * i];
  o[1] = c->X[2 * i + 1];
  o[2] = c->w[i];
}

/* common() (voronoi_triangulation_intersection.hpp:27-40) */
static int common_vertex(const pedge *f, const pedge *g) {
  if (f->a == g->a) return f->a;
  if (f->a == g->b) return f->a;
  return f->b;
}

/* inside() (voronoi_triangulation_intersection.hpp:125-166): is the vertex
 * a cap b strictly closer (power) to v than to w? */
static int ti_inside(const tictx *c, int64_t wv, const pedge *a, const pedge *b) {
  double V[3], Wp[3];
  wp(c, c->v, V);
  wp(c, wv, Wp);
  if (a->type == EDGE_T && b->type == EDGE_T) {
    int p = common_vertex(a, b);
    return orc_side1(V, Wp, c->t->x[p], c->t->y[p]) < 0;
  }
  if (a->type == EDGE_DT && b->type == EDGE_DT) {
    double U2[3], U3[3];
    wp(c, b->b, U2);
    wp(c, a->b, U3);
    return orc_side2(V, U2, U3, Wp) < 0;
  }
  int p, q;
  int64_t u;
  if (a->type == EDGE_T) {
    p = a->a; q = a->b; u = b->b;
  } else {
    p = b->a; q = b->b; u = a->b;
  }
  double U[3];
  wp(c, u, U);
  return orc_side3(V, U, c->t->x[p], c->t->y[p], c->t->x[q], c->t->y[q], Wp) < 0;
}

/* operator() (voronoi_triangulation_intersection.hpp:168-198) */
static int ti_clip(const tictx *c, const pedge *P, int n, int64_t wv, pedge *R) {
  if (n == 0) return 0;
  pedge L = {EDGE_DT, (int)c->v, (int)wv};
  int m = 0;
  int prev_in = ti_inside(c, wv, &P[n - 1], &P[0]);
  for (int i = 0; i < n; i++) {
    int ii = (i + 1) % n;
    int cur_in = ti_inside(c, wv, &P[i], &P[ii]);
    if (m + 2 > MAXPE) return -1;
    if (prev_in) {
      R[m++] = P[i];
      if (!cur_in) R[m++] = L;
    } else if (cur_in) {
      R[m++] = P[i];
    }
    prev_in = cur_in;
  }
  return m;
}

/* edge_to_line (voronoi_triangulation_intersection.hpp:98-110), as
 * a x + b y + c = 0 with CGAL's Line_2(p,q) and radical_axis conventions,
 * in coordinates relative to the origin (ox, oy).  The reference constructs
 * these lines in absolute coordinates (inexact constructions), where the
 * radical axis of two nearly coincident Diracs loses accuracy by
 * cancellation (|P|^2 - |Q|^2 with |P - Q| << |P|); relative to the Dirac
 * whose cell is being cut the same construction is accurate. */
static void edge_line(const tictx *c, const pedge *e, double ox, double oy, double *l) {
  if (e->type == EDGE_T) {
    double px = c->t->x[e->a] - ox, py = c->t->y[e->a] - oy;
    double qx = c->t->x[e->b] - ox, qy = c->t->y[e->b] - oy;
    l[0] = py - qy;
    l[1] = qx - px;
    l[2] = px * qy - py * qx;
  } else {
    double P[3], Q[3];
    wp(c, c->v, P);
    wp(c, e->b, Q);
    P[0] -= ox; P[1] -= oy;
    Q[0] -= ox; Q[1] -= oy;
    l[0] = 2.0 * (P[0] - Q[0]);
    l[1] = 2.0 * (P[1] - Q[1]);
    l[2] = -(P[0] * P[0] + P[1] * P[1]) + P[2] + (Q[0] * Q[0] + Q[1] * Q[1]) - Q[2];
  }
}
/* vertex_to_point / line_line_intersection (predicates.hpp:21-30), formed
 * relative to the Dirac v and translated back. */
static void vertex_point(const tictx *c, const pedge *e1, const pedge *e2, double *x,
                         double *y) {
  const double ox = c->X[2 * c->v], oy = c->X[2 * c->v + 1];
  double l1[3], l2[3];
  edge_line(c, e1, ox, oy, l1);
  edge_line(c, e2, ox, oy, l2);
  double den = l1[0] * l2[1] - l2[0] * l1[1];
  *x = ox + (l1[1] * l2[2] - l1[2] * l2[1]) / den;
  *y = oy + (l1[2] * l2[0] - l1[0] * l2[2]) / den;
}

/* ====================================================================== */
/* Driver: kantorovich (kantorovich.hpp:35-141) / first_moment (mode 1,   */
/* lloyd.hpp:30-69) / second_moment (mode 2, lloyd.hpp:71-123) / fragment */
/* polygons (mode 3, voronoi_triangulation_intersection.hpp:315-343).     */
/* ====================================================================== */
struct orc_result {
  int status;
  int64_t n;
  double fval;
  double *g, *mom, *mom2;
  int64_t nt, cap;
  int32_t *ti, *tj;
  double *tv;
  int64_t stats[6];
  /* mode 3: fragment f is Dirac fc[f] cut by triangle ft[f]; its vertices
     are fxy[2 * fs[f] ...] up to fs[f + 1] */
  int64_t nf, capf, nvx, capv;
  int32_t *fc, *ft;
  int64_t *fs;
  double *fxy;
};

static int push_frag(orc_result *r, int32_t v, int32_t f, int m, const double *px,
                     const double *py) {
  if (r->nf + 1 >= r->capf) {
    int64_t nc = r->capf ? 2 * r->capf : 1024;
    int32_t *a = realloc(r->fc, (size_t)nc * sizeof(int32_t));
    if (!a) return -1;
    r->fc = a;
    int32_t *b = realloc(r->ft, (size_t)nc * sizeof(int32_t));
    if (!b) return -1;
    r->ft = b;
    int64_t *c = realloc(r->fs, (size_t)(nc + 1) * sizeof(int64_t));
    if (!c) return -1;
    r->fs = c;
    r->capf = nc;
  }
  if (r->nvx + m > r->capv) {
    int64_t nc = r->capv ? 2 * r->capv : 4096;
    while (nc < r->nvx + m) nc *= 2;
    double *d = realloc(r->fxy, (size_t)(2 * nc) * sizeof(double));
    if (!d) return -1;
    r->fxy = d;
    r->capv = nc;
  }
  r->fc[r->nf] = v;
  r->ft[r->nf] = f;
  r->fs[r->nf] = r->nvx;
  for (int k = 0; k < m; k++) {
    r->fxy[2 * (r->nvx + k)] = px[k];
    r->fxy[2 * (r->nvx + k) + 1] = py[k];
  }
  r->nvx += m;
  r->nf++;
  r->fs[r->nf] = r->nvx;
  return 0;
}

static int push_trip(orc_result *r, int32_t i, int32_t j, double v) {
  if (r->nt == r->cap) {
    int64_t nc = r->cap ? 2 * r->cap : 1024;
    int32_t *a = realloc(r->ti, (size_t)nc * sizeof(int32_t));
    if (!a) return -1;
    r->ti = a;
    int32_t *b = realloc(r->tj, (size_t)nc * sizeof(int32_t));
    if (!b) return -1;
    r->tj = b;
    double *c = realloc(r->tv, (size_t)nc * sizeof(double));
    if (!c) return -1;
    r->tv = c;
    r->cap = nc;
  }
  r->ti[r->nt] = i;
  r->tj[r->nt] = j;
  r->tv[r->nt] = v;
  r->nt++;
  return 0;
}

orc_result *orc_run(int mode, int64_t nv, const double *vxy, const double *vval, int64_t nt,
                    const int32_t *tri, int64_t n, const double *X, const double *w) {
  orc_result *R = calloc(1, sizeof(orc_result));
  if (!R) return NULL;
  R->n = n;
  R->g = calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  R->mom = calloc((size_t)(2 * (n > 0 ? n : 1)), sizeof(double));
  R->mom2 = calloc((size_t)(3 * (n > 0 ? n : 1)), sizeof(double));
  g_npred = g_nexact = 0;
  if (nv <= 0 || nt <= 0 || n <= 0) return R;

  /* triangles: ccw orientation + Linear_function */
  otri *T = malloc((size_t)nt * sizeof(otri));
  double bx0 = INFINITY, by0 = INFINITY, bx1 = -INFINITY, by1 = -INFINITY;
  for (int64_t f = 0; f < nt; f++) {
    int32_t id[3] = {tri[3 * f], tri[3 * f + 1], tri[3 * f + 2]};
    double ar = tri_area(vxy[2 * id[0]], vxy[2 * id[0] + 1], vxy[2 * id[1]],
                         vxy[2 * id[1] + 1], vxy[2 * id[2]], vxy[2 * id[2] + 1]);
    if (ar < 0) {
      int32_t s = id[1];
      id[1] = id[2];
      id[2] = s;
    }
    double fv[3];
    for (int k = 0; k < 3; k++) {
      T[f].x[k] = vxy[2 * id[k]];
      T[f].y[k] = vxy[2 * id[k] + 1];
      fv[k] = vval[id[k]];
      bx0 = fmin(bx0, T[f].x[k]); bx1 = fmax(bx1, T[f].x[k]);
      by0 = fmin(by0, T[f].y[k]); by1 = fmax(by1, T[f].y[k]);
    }
    linear_function(&T[f], fv);
  }
  double diam = fmax(bx1 - bx0, by1 - by0);

  /* grid over the domain box and all sites */
  double gx0 = bx0, gy0 = by0, gx1 = bx1, gy1 = by1, W = -INFINITY;
  for (int64_t i = 0; i < n; i++) {
    gx0 = fmin(gx0, X[2 * i]); gx1 = fmax(gx1, X[2 * i]);
    gy0 = fmin(gy0, X[2 * i + 1]); gy1 = fmax(gy1, X[2 * i + 1]);
    W = fmax(W, w[i]);
  }
  double pad = 1e-9 * fmax(gx1 - gx0, gy1 - gy0) + 1e-300;
  gx0 -= pad; gy0 -= pad; gx1 += pad; gy1 += pad;
  double area = (gx1 - gx0) * (gy1 - gy0);
  double h = sqrt(area / fmax(1.0, (double)n / 2.0));
  int gx = (int)ceil((gx1 - gx0) / h), gy = (int)ceil((gy1 - gy0) / h);
  if (gx < 1) gx = 1;
  if (gy < 1) gy = 1;
  while ((double)gx * gy > 4e7) {
    h *= 1.5;
    gx = (int)ceil((gx1 - gx0) / h);
    gy = (int)ceil((gy1 - gy0) / h);
  }
  int nb = gx * gy;

  cellctx C;
  C.n = n; C.X = X; C.w = w; C.W = W;
  C.bx0 = bx0; C.by0 = by0; C.bx1 = bx1; C.by1 = by1;
  C.delta = 1e-9 * diam;
  C.sg.x0 = gx0; C.sg.y0 = gy0; C.sg.h = h; C.sg.gx = gx; C.sg.gy = gy;
  C.sg.start = calloc((size_t)nb + 1, sizeof(int32_t));
  C.sg.items = malloc((size_t)n * sizeof(int32_t));
  for (int64_t i = 0; i < n; i++)
    C.sg.start[bin_y(&C.sg, X[2 * i + 1]) * gx + bin_x(&C.sg, X[2 * i]) + 1]++;
  for (int b = 0; b < nb; b++) C.sg.start[b + 1] += C.sg.start[b];
  int32_t *cur = malloc((size_t)nb * sizeof(int32_t));
  memcpy(cur, C.sg.start, (size_t)nb * sizeof(int32_t));
  for (int64_t i = 0; i < n; i++) {
    int b = bin_y(&C.sg, X[2 * i + 1]) * gx + bin_x(&C.sg, X[2 * i]);
    C.sg.items[cur[b]++] = (int32_t)i;
  }
  /* triangle grid: same geometry, each triangle in every bin its box meets */
  ogrid tg = C.sg;
  tg.start = calloc((size_t)nb + 1, sizeof(int32_t));
  int64_t ntot = 0;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int b = 0; b < nb; b++) tg.start[b + 1] += tg.start[b];
      ntot = tg.start[nb];
      tg.items = malloc((size_t)(ntot > 0 ? ntot : 1) * sizeof(int32_t));
      memcpy(cur, tg.start, (size_t)nb * sizeof(int32_t));
    }
    for (int64_t f = 0; f < nt; f++) {
      double x0 = fmin(T[f].x[0], fmin(T[f].x[1], T[f].x[2]));
      double x1 = fmax(T[f].x[0], fmax(T[f].x[1], T[f].x[2]));
      double y0 = fmin(T[f].y[0], fmin(T[f].y[1], T[f].y[2]));
      double y1 = fmax(T[f].y[0], fmax(T[f].y[1], T[f].y[2]));
      for (int yy = bin_y(&tg, y0); yy <= bin_y(&tg, y1); yy++)
        for (int xx = bin_x(&tg, x0); xx <= bin_x(&tg, x1); xx++) {
          int b = yy * gx + xx;
          if (pass == 0)
            tg.start[b + 1]++;
          else
            tg.items[cur[b]++] = (int32_t)f;
        }
    }
  }
  free(cur);

  int64_t *stamp = malloc((size_t)nt * sizeof(int64_t));
  for (int64_t f = 0; f < nt; f++) stamp[f] = -1;
  opoly *P = malloc(sizeof(opoly));
  int32_t *nbr = malloc(sizeof(int32_t) * 8192);
  pedge *E1 = malloc(sizeof(pedge) * MAXPE), *E2 = malloc(sizeof(pedge) * MAXPE);
  double *px = malloc(sizeof(double) * MAXPE), *py = malloc(sizeof(double) * MAXPE);
  double fval = 0.0;

  for (int64_t v = 0; v < n && R->status == 0; v++) {
    double bb[4] = {0.0, 0.0, 0.0, 0.0};
    int nn = power_cell(&C, v, P, nbr, 8192, bb);
    if (nn < 0) {
      R->status = 2;
      break;
    }
    if (P->n == 0) continue;
    R->stats[4]++;
    R->stats[5] += nn;
    tictx tc;
    tc.X = X; tc.w = w; tc.v = v;
    const double yvx = X[2 * v], yvy = X[2 * v + 1];
    for (int yy = bin_y(&tg, bb[1]); yy <= bin_y(&tg, bb[3]); yy++)
      for (int xx = bin_x(&tg, bb[0]); xx <= bin_x(&tg, bb[2]); xx++) {
        int b = yy * gx + xx;
        for (int32_t q = tg.start[b]; q < tg.start[b + 1]; q++) {
          int32_t f = tg.items[q];
          if (stamp[f] == v) continue;
          stamp[f] = v;
          const otri *t = &T[f];
          double x0 = fmin(t->x[0], fmin(t->x[1], t->x[2]));
          double x1 = fmax(t->x[0], fmax(t->x[1], t->x[2]));
          double y0 = fmin(t->y[0], fmin(t->y[1], t->y[2]));
          double y1 = fmax(t->y[0], fmax(t->y[1], t->y[2]));
          if (x1 < bb[0] || x0 > bb[2] || y1 < bb[1] || y0 > bb[3]) continue;
          R->stats[0]++;
          tc.t = t;
          /* triangle as a polygon of 3 EDGE_T edges
           * (voronoi_triangulation_intersection.hpp:259-263) */
          int m = 3;
          E1[0].type = EDGE_T; E1[0].a = 0; E1[0].b = 1;
          E1[1].type = EDGE_T; E1[1].a = 1; E1[1].b = 2;
          E1[2].type = EDGE_T; E1[2].a = 2; E1[2].b = 0;
          pedge *cp = E1, *np = E2;
          for (int k = 0; k < nn && m > 0; k++) {
            m = ti_clip(&tc, cp, m, nbr[k], np);
            if (m < 0) {
              R->status = 3;
              break;
            }
            pedge *s = cp;
            cp = np;
            np = s;
          }
          if (m < 3) continue;
          R->stats[1]++;
          /* polygon vertices (kantorovich.hpp:93-102) */
          for (int k = 0; k < m; k++) {
            int kk = (k == 0) ? m - 1 : k - 1;
            vertex_point(&tc, &cp[k], &cp[kk], &px[k], &py[k]);
          }
          if (mode == 3) {
            /* the polygon handed to the traversal's callback
               (voronoi_triangulation_intersection.hpp:329-338); fragments of
               zero area (a cell touching the triangle along an edge or at a
               point) are left out, with the same scale-aware threshold on
               coordinates local to the Dirac as the integrals use */
            double ar2 = 0.0, xlo = INFINITY, xhi = -INFINITY, ylo = INFINITY, yhi = -INFINITY;
            double cm = 0.0;
            for (int k = 0; k < m; k++) {
              int k1 = (k + 1) % m;
              double ax = px[k] - yvx, ay = py[k] - yvy;
              double bx = px[k1] - yvx, by = py[k1] - yvy;
              ar2 += ax * by - bx * ay;
              xlo = fmin(xlo, ax); xhi = fmax(xhi, ax);
              ylo = fmin(ylo, ay); yhi = fmax(yhi, ay);
              cm = fmax(cm, fmax(fabs(ax), fabs(ay)));
            }
            if (!(0.5 * ar2 > 1e-13 * ((xhi - xlo) + (yhi - ylo)) * cm)) continue;
            if (push_frag(R, (int32_t)v, (int32_t)f, m, px, py)) {
              R->status = 4;
              break;
            }
          } else if (mode == 0) {
            /* Hessian triplets (kantorovich.hpp:109-122) */
            for (int k = 0; k < m; k++) {
              if (cp[k].type != EDGE_DT) continue;
              int64_t wv = cp[k].b;
              int k1 = (k + 1) % m;
              double ex = px[k1] - px[k], ey = py[k1] - py[k];
              double r = sqrt(ex * ex + ey * ey) *
                         rho_at(t, (px[k] + px[k1]) / 2.0, (py[k] + py[k1]) / 2.0);
              double dx = X[2 * v] - X[2 * wv], dy = X[2 * v + 1] - X[2 * wv + 1];
              double d = 2.0 * sqrt(dx * dx + dy * dy);
              if (push_trip(R, (int32_t)v, (int32_t)wv, -r / d) ||
                  push_trip(R, (int32_t)v, (int32_t)v, +r / d)) {
                R->status = 4;
                break;
              }
            }
            /* value and gradient (kantorovich.hpp:124-135) */
            double warea = 0.0, intg = 0.0;
            kctx kc = {t, yvx, yvy};
            for (int k = 1; k < m - 1; k++) {
              warea += integrate_centroid(px[0], py[0], px[k], py[k], px[k + 1], py[k + 1], t);
              double r1;
              integrate_ac6(px[0], py[0], px[k], py[k], px[k + 1], py[k + 1], f_intg, &kc, 1,
                            &r1);
              intg += r1;
            }
            fval = fval + warea * w[v] - intg;
            R->g[v] += warea;
          } else if (mode == 2) {
            /* second_moment (lloyd.hpp:71-123): all six integrals by the
               Albrecht-Collatz rule (integrate_3, quadrature.hpp:133-143) */
            kctx kc = {t, yvx, yvy};
            for (int k = 1; k < m - 1; k++) {
              double r6[6];
              integrate_ac6(px[0], py[0], px[k], py[k], px[k + 1], py[k + 1], f_moment2, &kc,
                            6, r6);
              R->g[v] += r6[0];
              R->mom[2 * v] += r6[1];
              R->mom[2 * v + 1] += r6[2];
              R->mom2[3 * v] += r6[3];
              R->mom2[3 * v + 1] += r6[4];
              R->mom2[3 * v + 2] += r6[5];
            }
          } else {
            /* first_moment (lloyd.hpp:30-69) */
            kctx kc = {t, yvx, yvy};
            for (int k = 1; k < m - 1; k++) {
              double r3[3];
              integrate_ac6(px[0], py[0], px[k], py[k], px[k + 1], py[k + 1], f_moment, &kc,
                            3, r3);
              R->g[v] += r3[0];
              R->mom[2 * v] += r3[1];
              R->mom[2 * v + 1] += r3[2];
            }
          }
        }
      }
  }
  R->fval = fval;
  R->stats[2] = g_npred;
  R->stats[3] = g_nexact;
  free(stamp);
  free(P);
  free(nbr);
  free(E1);
  free(E2);
  free(px);
  free(py);
  free(T);
  grid_free(&C.sg);
  grid_free(&tg);
  return R;
}

int orc_status(const orc_result *r) { return r ? r->status : -1; }
int64_t orc_ntriplets(const orc_result *r) { return r ? r->nt : 0; }
void orc_copy(const orc_result *r, double *fval, double *g, double *mom, int32_t *ti,
              int32_t *tj, double *tv, int64_t *stats) {
  if (fval) *fval = r->fval;
  if (g && r->n > 0) memcpy(g, r->g, (size_t)r->n * sizeof(double));
  if (mom && r->n > 0) memcpy(mom, r->mom, (size_t)(2 * r->n) * sizeof(double));
  if (stats) memcpy(stats, r->stats, sizeof r->stats);
  if (ti && r->nt) memcpy(ti, r->ti, (size_t)r->nt * sizeof(int32_t));
  if (tj && r->nt) memcpy(tj, r->tj, (size_t)r->nt * sizeof(int32_t));
  if (tv && r->nt) memcpy(tv, r->tv, (size_t)r->nt * sizeof(double));
}
void orc_copy_mom2(const orc_result *r, double *mom2) {
  if (mom2 && r->n > 0) memcpy(mom2, r->mom2, (size_t)(3 * r->n) * sizeof(double));
}
int64_t orc_nfrag(const orc_result *r, int64_t *nvert) {
  if (nvert) *nvert = r ? r->nvx : 0;
  return r ? r->nf : 0;
}
void orc_copy_frags(const orc_result *r, int32_t *cell, int32_t *tri, int64_t *start,
                    double *xy) {
  if (!r || r->nf == 0) {
    if (start) start[0] = 0;
    return;
  }
  memcpy(cell, r->fc, (size_t)r->nf * sizeof(int32_t));
  memcpy(tri, r->ft, (size_t)r->nf * sizeof(int32_t));
  memcpy(start, r->fs, (size_t)(r->nf + 1) * sizeof(int64_t));
  memcpy(xy, r->fxy, (size_t)(2 * r->nvx) * sizeof(double));
}
void orc_free(orc_result *r) {
  if (!r) return;
  free(r->fc);
  free(r->ft);
  free(r->fs);
  free(r->fxy);
  free(r->g);
  free(r->mom);
  free(r->mom2);
  free(r->ti);
  free(r->tj);
  free(r->tv);
  free(r);
}

/* ====================================================================== */
/* draw_laguerre_diagram (rasterization.hpp:491-547): fragments from mode  */
/* 3, each coloured by colors[Dirac] times the mean of rho over its        */
/* vertices (:528-534) and spread over pixels by exact coverage: the       */
/* polygon clipped by the pixel square (Sutherland-Hodgman, four axis      */
/* lines), area by the shoelace formula.  The reference's scan converter   */
/* (rasterize_convex_polygon, :403-470) computes the same coverage after   */
/* nudging vertices within 1e-8 of an integer by 1e-6 pixel.              */
/* ====================================================================== */
static int clip_axis(const double *ix, const double *iy, int n, int axis, double c, int keep_ge,
                     double *ox, double *oy) {
  int m = 0;
  for (int i = 0; i < n; i++) {
    int j = (i + 1) % n;
    double a = axis ? iy[i] : ix[i], b = axis ? iy[j] : ix[j];
    int ina = keep_ge ? a >= c : a <= c, inb = keep_ge ? b >= c : b <= c;
    if (ina) {
      ox[m] = ix[i];
      oy[m] = iy[i];
      m++;
    }
    if (ina != inb) {
      double t = (c - a) / (b - a);
      ox[m] = ix[i] + t * (ix[j] - ix[i]);
      oy[m] = iy[i] + t * (iy[j] - iy[i]);
      if (axis) oy[m] = c; else ox[m] = c;
      m++;
    }
  }
  return m;
}

int orc_draw_laguerre(int64_t nv, const double *vxy, const double *vval, int64_t nt,
                      const int32_t *tri, int64_t n, const double *X, const double *w,
                      const double *colors, int nc, double x0, double y0, double x1,
                      double y1, int W, int H, double *img) {
  orc_result *R = orc_run(3, nv, vxy, vval, nt, tri, n, X, w);
  if (!R) return -1;
  if (R->status) {
    int s = R->status;
    orc_free(R);
    return s;
  }
  const double sx = (double)W / (x1 - x0), sy = (double)H / (y1 - y0);
  double ax[128], ay[128], bx[128], by[128];
  for (int64_t f = 0; f < R->nf; f++) {
    int64_t s0 = R->fs[f];
    int m = (int)(R->fs[f + 1] - s0);
    if (m > 100) m = 100; /* longer fragments do not arise in the tests */
    /* density: the triangle's linear function (functions.hpp:23-80) */
    otri t;
    int32_t id[3] = {tri[3 * R->ft[f]], tri[3 * R->ft[f] + 1], tri[3 * R->ft[f] + 2]};
    double fv[3];
    for (int k = 0; k < 3; k++) {
      t.x[k] = vxy[2 * id[k]];
      t.y[k] = vxy[2 * id[k] + 1];
      fv[k] = vval[id[k]];
    }
    linear_function(&t, fv);
    double ff = 0.0, px[128], py[128];
    double lo_x = INFINITY, hi_x = -INFINITY, lo_y = INFINITY, hi_y = -INFINITY;
    for (int k = 0; k < m; k++) {
      double x = R->fxy[2 * (s0 + k)], y = R->fxy[2 * (s0 + k) + 1];
      ff += rho_at(&t, x, y);
      px[k] = sx * (x - x0);
      py[k] = sy * (y - y0);
      lo_x = fmin(lo_x, px[k]); hi_x = fmax(hi_x, px[k]);
      lo_y = fmin(lo_y, py[k]); hi_y = fmax(hi_y, py[k]);
    }
    ff /= (double)m;
    const double *col = colors + (size_t)R->fc[f] * nc;
    int i0 = (int)floor(lo_x), i1 = (int)floor(hi_x), j0 = (int)floor(lo_y), j1 = (int)floor(hi_y);
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (i1 > W - 1) i1 = W - 1;
    if (j1 > H - 1) j1 = H - 1;
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        int q = clip_axis(px, py, m, 0, (double)i, 1, ax, ay);
        q = clip_axis(ax, ay, q, 0, (double)(i + 1), 0, bx, by);
        q = clip_axis(bx, by, q, 1, (double)j, 1, ax, ay);
        q = clip_axis(ax, ay, q, 1, (double)(j + 1), 0, bx, by);
        double a2 = 0.0;
        for (int k = 0; k < q; k++) {
          int k1 = (k + 1) % q;
          a2 += bx[k] * by[k1] - bx[k1] * by[k];
        }
        double cover = 0.5 * a2;
        if (!(cover > 0.0)) continue;
        for (int c = 0; c < nc; c++) img[((size_t)j * W + i) * nc + c] += cover * ff * col[c];
      }
  }
  orc_free(R);
  return 0;
}
