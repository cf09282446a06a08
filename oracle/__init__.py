"""CPU oracle (test infrastructure only): FP64 restatement of the reference
hot path.  Imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py."""
from .oracle import (OracleStats, SolveStats, ac6_poly, centroid_lin, draw_laguerre_diagram,
                     first_moment, fragments,
                     kantorovich, kantorovich_triplets, lloyd, ot_solve, second_moment, side1, side1_exact,
                     side2, side2_exact, side3, side3_exact, solve_laplacian_matrix,
                     voronoi_polygon_intersection)
