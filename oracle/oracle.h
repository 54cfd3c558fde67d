/*
 * CPU oracle for the epsilon-neighbor-graph path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (rnng_amd/, include/,
 * the C-ABI library) links, loads or calls this code.  It is the checker:
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it
 * to compute the expected answer on the CPU.
 *
 * Every function restates the reference algorithm in plain C and cites the
 * reference file:line it follows.  Build flags (see Makefile) pin the
 * floating-point semantics the reference's own Release build has: x86-64
 * baseline code (no -march, so no FMA) and -ffp-contract=off.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_l2sq / orc_l2        known-answer vectors with exact integer results,
 *                            a double-precision bound, a planted-FMA check
 *                            (tests/test_oracle.py)
 *   orc_sq_threshold         exhaustive neighbourhood check of the defining
 *                            property sqrtf(v) <= eps  <=>  v <= T(eps)
 *   orc_l2_eps_pairs         brute force in numpy with the same per-pair
 *                            sequential order, symmetry, eps-monotonicity
 *   orc_edit_distance        textbook known answers (kitten/sitting = 3, ...),
 *                            closed forms (empty string, identical strings,
 *                            pure substitution = Hamming), a recursive
 *                            definition on tiny inputs
 *   orc_voronoi_assign       brute-force argmin with the later-wins tie rule,
 *                            planted equal-distance ties
 *   orc_fps                  a numpy restatement on small inputs (farthest
 *                            point by explicit distance matrix), the
 *                            by-product assignment equal to orc_voronoi_assign
 *                            of the chosen centers
 *   orc_ghosts               the covering property: the union of cell-local
 *                            pairs equals the brute-force graph
 *   orc_binom / orc_simplex_id  closed-form C(n,k), the combinatorial
 *                            number system's bijectivity on tiny n
 *   orc_rips_filtration      brute-force triple enumeration, K_n clique
 *                            counts C(n,k)
 */
#ifndef RNNG_ORACLE_H_
#define RNNG_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference-order squared L2 distance: drivers/systolic.cpp:253-269
 * (same functor in drivers/landmark.cpp:335-352, drivers/rips.cpp:326-342).
 * delta = fl(p_k - q_k); val = fl(val + fl(delta*delta)), k = 0..d-1. */
float orc_l2sq(const float *p, const float *q, int64_t d);

/* sqrtf(orc_l2sq(p, q, d)): the value the reference's L2Distance returns. */
float orc_l2(const float *p, const float *q, int64_t d);

/* Largest float T >= 0 with sqrtf(T) <= eps (-1 if none, e.g. eps < 0 or NaN).
 * sqrtf is monotone, so sqrtf(v) <= eps  <=>  v <= T for every float v >= 0.
 * The reference tests dist <= radius with radius held as float
 * (drivers/systolic.cpp:225; include/search.hpp:244,399). */
float orc_sq_threshold(float eps);

/* Brute-force epsilon range (include/search.hpp:389-407 semantics) over a
 * self join (self_join=1, X == Y, pairs i < j) or a bipartite join
 * (self_join=0, all (i, j) with i in X, j in Y).  A pair is an edge iff
 * sqrtf(orc_l2sq(x_i, y_j)) <= eps.  Output is row-major (i ascending, then j
 * ascending), ij[2*e] = i, ij[2*e+1] = j, w[e] = the float distance.  At most
 * cap edges are written; the return value is the total count. */
int64_t orc_l2_eps_pairs(const float *X, int64_t nx, const float *Y, int64_t ny,
                         int64_t d, float eps, int self_join, int nthreads,
                         int64_t *ij, float *w, int64_t cap);

/* Levenshtein distance with unit costs, two-row dynamic programme in int64:
 * drivers/systolic.cpp:271-299 (same in landmark.cpp:354-383, rips.cpp:344-372). */
int64_t orc_edit_distance(const char *s, int64_t m, const char *t, int64_t n);

/* Brute-force edit-distance epsilon graph over strings given as a CSR
 * (atoms, offsets[n+1]).  Pairs i < j with (float)dist <= eps.  The length
 * difference |m - n| is a lower bound of the distance and skips pairs whose
 * lengths alone exceed eps; it never changes the result.  Output row-major,
 * dist[e] = the integer distance. */
int64_t orc_edit_eps_pairs(const char *atoms, const int64_t *offsets, int64_t n,
                           float eps, int nthreads, int64_t *ij, int32_t *dist,
                           int64_t cap);

/* Nearest-center assignment, include/point.hpp:535-559: for each point,
 * best = FLT_MAX, cell = 0; for c = 0..m-1, dist = L2(center_c, point);
 * if dist <= best then (best, cell) = (dist, c).  Ties go to the later center. */
void orc_voronoi_assign(const float *pts, int64_t n, const float *centers, int64_t m,
                        int64_t d, int nthreads, int64_t *cell, float *dist);

/* Farthest-point landmarks; (cell, dist) = Voronoi assignment of the chosen
 * centers in selection order. */
void orc_fps(const float *pts, int64_t n, int64_t d, int64_t m, int64_t first, int nthreads,
             int64_t *centers, int64_t *cell, float *dist);

/* Ghost criterion of include/point.hpp:894-918 (single rank): point q with
 * own-center distance delta_q becomes a ghost of cell c != cell(q) iff c has
 * at least one point, L2(center_c, q) <= fl(delta_q + 2*eps) and some point p
 * of cell c has L2(p, q) <= eps.  cell[] and delta[] come from
 * orc_voronoi_assign; centers[] are the m center coordinates.  Output pairs
 * (c, q) sorted by c then q: out[2*g] = c, out[2*g+1] = q.  Returns the count. */
int64_t orc_ghosts(const float *pts, int64_t n, int64_t d, const int64_t *cell,
                   const float *delta, const float *centers, int64_t m, float eps,
                   int nthreads, int64_t *out, int64_t cap);

/* Binomial coefficient with the reference's semantics (include/utils.hpp:116-144):
 * 0 if k > n, else n(n-1)...(n-k+1)/k!.  Computed exactly (128-bit) so values
 * that fit int64 are returned exactly. */
int64_t orc_binom(int64_t n, int64_t k);

/* Simplex id, include/simplex.hpp:5-19: sort the k vertices ascending,
 * uid = sum_i C(v_i, i+1), id = uid | ((k-1) << 60). */
int64_t orc_simplex_id(const int64_t *verts, int64_t k);

/* Vietoris-Rips filtration up to maxdim (include/point.hpp:1134-1267 plus the
 * merge in include/simplex.hpp:130-196): every clique of the epsilon graph
 * with at most maxdim+1 vertices; value 0 for vertices, the float L2 distance
 * for edges, the max over facets for higher simplices.  Sorted by
 * (value, id), unique by id.  Returns the count. */
int64_t orc_rips_filtration(const float *pts, int64_t n, int64_t d, float eps,
                            int64_t maxdim, int nthreads, float *value, int64_t *id,
                            int64_t cap);

#ifdef __cplusplus
}
#endif

#endif
