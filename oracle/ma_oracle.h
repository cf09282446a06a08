/*
 * CPU oracle for the power-diagram x triangulation hot path of MongeAmpere++.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU engine is
 * compared against; nothing in the product links, loads or calls it.
 *
 * It restates, in plain C99 / FP64, the reference algorithm:
 *   include/MA/kantorovich.hpp:35-141          (value, masses, Hessian triplets)
 *   include/MA/voronoi_triangulation_intersection.hpp:42-199  (Tri_intersector)
 *   include/MA/predicates.hpp:21-168           (Side1/Side2/Side3, exact signs)
 *   include/MA/quadrature.hpp:23-42,69-85,122-143 (centroid, AC6, segment rules)
 *   include/MA/functions.hpp:23-80             (barycentric / Linear_function)
 *   include/MA/lloyd.hpp:30-69                 (first_moment)
 *   include/MA/lloyd.hpp:71-123                (second_moment)
 *   include/MA/voronoi_triangulation_intersection.hpp:315-343 (fragment polygons)
 *   include/MA/rasterization.hpp:403-547       (draw_laguerre_diagram)
 * The CGAL regular triangulation (kantorovich.hpp:64-72) is replaced by a
 * certified FP64 power-cell search that returns a superset of each Dirac's
 * power neighbours; clipping by a superset gives the same fragments.
 */
#ifndef MA_ORACLE_H
#define MA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_result orc_result;

/* mode 0: kantorovich (fval, g, Hessian triplets); mode 1: first_moment
 * (masses, first moments int rho*x, int rho*y per cell); mode 2:
 * second_moment; mode 3: the polygon of every fragment of positive area.
 * X is n x 2 row-major (x0,y0,x1,y1,...); tri is nt x 3 vertex indices. */
orc_result *orc_run(int mode, int64_t nv, const double *vxy, const double *vval,
                    int64_t nt, const int32_t *tri, int64_t n, const double *X,
                    const double *w);
int orc_status(const orc_result *r);
int64_t orc_ntriplets(const orc_result *r);
/* stats: [0]=#pairs clipped, [1]=#non-empty fragments, [2]=#predicates,
 *        [3]=#predicates decided by the exact stage, [4]=#cells, [5]=sum of
 *        neighbour-set sizes */
void orc_copy(const orc_result *r, double *fval, double *g, double *mom,
              int32_t *ti, int32_t *tj, double *tv, int64_t *stats);
void orc_free(orc_result *r);
void orc_copy_mom2(const orc_result *r, double *mom2);
/* mode 3 results: fragment count (and total vertex count); fragment f is
 * Dirac cell[f] cut by triangle tri[f], vertices xy[2*start[f] .. 2*start[f+1])
 * counter-clockwise.  start has nfrag + 1 entries. */
int64_t orc_nfrag(const orc_result *r, int64_t *nvert);
void orc_copy_frags(const orc_result *r, int32_t *cell, int32_t *tri, int64_t *start,
                    double *xy);
/* draw_laguerre_diagram (rasterization.hpp:491-547): every fragment, mapped
 * from the box [x0,x1]x[y0,y1] to a W x H pixel grid, adds
 * coverage(pixel) * mean of rho over its vertices * colors[cell*nc + c] to
 * img[(j*W + i)*nc + c], coverage = exact area of fragment cap pixel square
 * (pixel (i,j) = [i,i+1]x[j,j+1]); pixels outside the grid are dropped.
 * Returns the oracle status (0 = ok). */
int orc_draw_laguerre(int64_t nv, const double *vxy, const double *vval, int64_t nt,
                      const int32_t *tri, int64_t n, const double *X, const double *w,
                      const double *colors, int nc, double x0, double y0, double x1,
                      double y1, int W, int H, double *img);

/* Exact sign predicates (exposed for tests).  Each returns the sign of
 * pi_v(x) - pi_w(x) at the implied point x, pi_u(x) = |x-u|^2 - u_w.
 * side1: x = (px,py); side2: x = power vertex of (v,u2,u3);
 * side3: x = line(p,q) cap radical_axis(v,u).  v,w,... are (x,y,weight). */
int orc_side1(const double *v, const double *w, double px, double py);
int orc_side2(const double *v, const double *u2, const double *u3, const double *w);
int orc_side3(const double *v, const double *u, double px, double py, double qx,
              double qy, const double *w);
/* Same predicates evaluated only with the exact (expansion) stage. */
int orc_side1_exact(const double *v, const double *w, double px, double py);
int orc_side2_exact(const double *v, const double *u2, const double *u3, const double *w);
int orc_side3_exact(const double *v, const double *u, double px, double py, double qx,
                    double qy, const double *w);

/* Quadrature restatements (quadrature.hpp): integrate a cubic polynomial
 * given by its 10 monomial coefficients c[i] of x^a y^b, (a,b) in the order
 * 1,x,y,x2,xy,y2,x3,x2y,xy2,y3, over triangle (a,b,c) with the
 * Albrecht-Collatz rule; and a linear function with the centroid rule. */
double orc_ac6_poly(const double *tri6, const double *c10);
double orc_centroid_lin(const double *tri6, double a, double b, double c);

#ifdef __cplusplus
}
#endif
#endif
