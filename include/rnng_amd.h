/*
 * rnng_amd: C ABI of the MI355X (gfx950) epsilon-neighbor-graph library.
 *
 * Plain pointers and sizes only.  Every function returns an int status:
 * RNNG_OK (0) on success, a negative RNNG_ERR_* code on failure, with a
 * message available from rnng_last_error() on the calling thread.  No C++
 * exception crosses this boundary.
 *
 * Semantics follow the reference (gabe-raulet/RNNG):
 *   - Index = int64, Real = float (include/utils.h:24-25).
 *   - A pair (p, q) is an edge iff L2(p, q) <= eps, where L2 is the float
 *     sqrtf of the sequential FP32 sum of squared differences over dims
 *     0..d-1, not fused (drivers/systolic.cpp:253-269), and eps is a float
 *     (drivers/systolic.cpp:225; include/search.hpp:244,399).
 *   - The returned weight is exactly that float distance.
 *
 * Calls are synchronous from the host's point of view (results are complete
 * when the call returns), as the reference's functor-before-return contract
 * requires (include/search.hpp:223-264).
 */
#ifndef RNNG_AMD_H_
#define RNNG_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RNNG_OK            0
#define RNNG_ERR_ARG      -1   /* invalid argument (null pointer, bad size) */
#define RNNG_ERR_HIP      -2   /* HIP runtime error (message has the details) */
#define RNNG_ERR_CAPACITY -3   /* caller's output buffer too small; *needed is set */
#define RNNG_ERR_NODEV    -4   /* no usable gfx950 device */
#define RNNG_ERR_FORMAT   -5   /* malformed input file */
#define RNNG_ERR_UNSUPPORTED -6 /* shape the requested mode cannot handle */

/* Distance-stage algorithm. */
#define RNNG_MODE_PLAIN 0   /* FP32 tiled kernel in the reference's order */
#define RNNG_MODE_MFMA  1   /* MFMA candidate filter (BF16, or exact int8 when every coordinate is an
                               integer and each dimension spans <= 255) + FP32 recheck + FP64 band
                               check; d > 144 or coordinates no filter can bound (NaN, infinities,
                               d * range^2 > 1e36) run the PLAIN kernel (stats: plain_fallbacks) */
#define RNNG_MODE_GRID  2   /* self join, d <= 3: uniform grid of cells of side >= eps; only
                               pairs in adjacent cells are evaluated, each in the reference's
                               FP32 order (same edges as PLAIN); pair_evals counts them */

/* Counters and device times of one call.  Times are HIP-event times on the
 * library's stream, in milliseconds. */
typedef struct rnng_stats {
    int64_t pair_evals;    /* unordered pairs covered by the distance stage */
    int64_t candidates;    /* filter survivors (MFMA mode; 0 in plain mode) */
    int64_t edges;         /* undirected edges found (i < j for a self join) */
    int64_t band_pairs;    /* candidates with |d - eps| <= 1e-5 * eps */
    int64_t band_flips;    /* band pairs whose FP64 decision differs from FP32's */
    int64_t refills;       /* count-then-fill reruns after an undersized buffer */
    double prep_ms;        /* BF16 fragment preparation (MFMA mode) */
    double filter_ms;      /* distance kernel (MFMA filter or plain kernel) */
    double recheck_ms;     /* FP32 recheck + band check (MFMA mode) */
    double total_ms;       /* all device work of the call */
    int64_t ghost_tests;      /* Voronoi: exact center tests of cross-cell edges */
    int64_t ghost_candidates; /* Voronoi: center tests passed (with repeats) */
    int64_t ghosts;           /* Voronoi: distinct (cell, point) ghosts */
    double ghost_ms;          /* Voronoi: cell adjacency, layout, cost-model features */
    int64_t filter_passes;    /* Voronoi: filter launches (2 for count-then-fill) */
    int64_t predicted_candidates; /* Voronoi: cost model's candidate estimate (0 without) */
    int64_t predicted_edges;      /* Voronoi: cost model's edge estimate (0 without) */
    int64_t int8_filter;          /* MFMA builds that ran the exact int8 filter (integer data) */
    int64_t plain_fallbacks;      /* MFMA builds run by the plain kernel: coordinates no filter can
                                     bound (NaN, infinities, d * range^2 > 1e36) */
    double call_ms;               /* host wall time of the build calls (graph and Voronoi builds) */
    double host_ms;               /* Voronoi: host time spent on the cell layout and work lists */
    double kernel_ms;             /* one stage kernel's device time: the banded / bit-parallel verify of
                                     an edit build, the triangle fill of a Rips build */
    int64_t kernel_units;         /* its work: text columns the verify ran (pairs inside the length
                                     window), triangles the fill wrote */
} rnng_stats;

/* Opaque handles. */
typedef struct rnng_dataset_s *rnng_dataset_t;   /* device-resident point set */
typedef struct rnng_graph_s *rnng_graph_t;       /* device-resident edge list */

/* Library version (major*10000 + minor*100 + patch). */
int rnng_version(void);

/* Message of the last failure on this thread ("" if none). */
const char *rnng_last_error(void);

/* Select and initialise the GPU (creates the library's stream).  Called
 * implicitly with device 0 by every other entry point if not called first. */
int rnng_init(int device);

/* ------------------------------------------------------------------------ */
/* Device-resident point sets                                               */
/* ------------------------------------------------------------------------ */

/* Upload n points of d floats (row-major) to HBM.  The PointContainer
 * layout of the reference (include/point.h:34-100) with fixed dimension:
 * point i is points[i*d .. i*d+d). */
int rnng_dataset_create_f32(const float *points, int64_t n, int32_t d, rnng_dataset_t *out);

/* Overwrites a dataset created by rnng_dataset_create_f32 (or read from a
 * file) with n points of its dimension, n no more than it was created with: one
 * device buffer reused for changing query batches, with no allocation (the
 * reference reuses its receive buffers the same way, PointContainer::sendrecv,
 * include/point.hpp:369-417). */
int rnng_dataset_upload_f32(rnng_dataset_t ds, const float *points, int64_t n);

/* Wrap points that already live in device memory (not copied, not freed). */
int rnng_dataset_wrap_device_f32(const float *dev_points, int64_t n, int32_t d,
                                 rnng_dataset_t *out);

int rnng_dataset_destroy(rnng_dataset_t ds);

/* Points and dimension of a dataset. */
int rnng_dataset_shape(rnng_dataset_t ds, int64_t *n, int32_t *d);

/* Copy a dataset's n*d floats back to the host. */
int rnng_dataset_copy_f32(rnng_dataset_t ds, float *out);

/* Point files straight into HBM: .fvecs (include/point.hpp:112-168), .fbin
 * and .u8bin (dataset_io.py:48-70; u8 values become floats), chosen by
 * extension as dataset_io.py's read_file does, points [start, start + count)
 * (count < 0: to the end).  Chunks are read into pinned host buffers and
 * uploaded asynchronously while the next chunk is read; u8bin travels as
 * bytes and fvecs as records, converted on the device.  Errors:
 * RNNG_ERR_FORMAT for an unknown extension, a header that does not match the
 * file size or mixed fvecs dimensions; RNNG_ERR_ARG for a range beyond the
 * file. */
int rnng_dataset_read_file(const char *path, int64_t start, int64_t count, rnng_dataset_t *out);

/* Header of a point file: points, dimension, kind (0 float32, 1 uint8). */
int rnng_file_info(const char *path, int64_t *n, int32_t *d, int32_t *kind);

/* ------------------------------------------------------------------------ */
/* L2 epsilon graph                                                          */
/* ------------------------------------------------------------------------ */

/* All pairs i < j of ds with L2 <= eps (the systolic driver's all-pairs
 * query, drivers/systolic.cpp:123-147, over one GPU's whole point set).
 * The edge list stays on the device in *out until copied. */
int rnng_graph_build(rnng_dataset_t ds, float eps, int32_t mode, rnng_graph_t *out,
                     rnng_stats *stats);

/* All pairs (i, j), i in a, j in b, with L2 <= eps (block x block, cell x
 * ghosts: include/search.hpp:266-340 over two point sets). */
int rnng_graph_build_bipartite(rnng_dataset_t a, rnng_dataset_t b, float eps, int32_t mode,
                               rnng_graph_t *out, rnng_stats *stats);

int rnng_graph_num_edges(rnng_graph_t g, int64_t *n_edges);

/* Copy up to cap edges: ij[2e] = i, ij[2e+1] = j, w[e] = float distance
 * (either may be NULL).  sorted != 0 orders them by (i, j). */
int rnng_graph_copy(rnng_graph_t g, int64_t *ij, float *w, int64_t cap, int32_t sorted);

/* The edge list the reference's drivers hold after Graph::redistribute_edges
 * on one rank (include/graph.hpp:59-116): for a self-join graph over
 * n_vertices points, (i, j, w) and (j, i, w) for every edge plus (v, v, 0) for
 * every vertex (the query of a point finds itself), sorted by (u, v).
 * *needed = 2 E + n_vertices; with cap < *needed and NULL outputs only the
 * size is reported (RNNG_OK), with non-NULL outputs RNNG_ERR_CAPACITY. */
/* The same symmetric list as CSR, built on the device: rowptr[0..n_vertices]
 * (row v = positions rowptr[v] .. rowptr[v+1]), col/w in (row, column)
 * order, with the (v, v, 0) loops if self_loops != 0.  *needed = 2 E (+ n);
 * rowptr may be requested alone; col/w need cap >= *needed
 * (RNNG_ERR_CAPACITY otherwise).  An edge naming a vertex >= n_vertices is
 * RNNG_ERR_ARG.  No reference counterpart: the reference keeps the edge list
 * (include/graph.h:7-27); this is its symmetrised form for CSR consumers. */
int rnng_graph_copy_csr(rnng_graph_t g, int64_t n_vertices, int32_t self_loops, int64_t *rowptr, int32_t *col,
                        float *w, int64_t cap, int64_t *needed);

int rnng_graph_copy_directed(rnng_graph_t g, int64_t n_vertices, int64_t *uv, float *w, int64_t cap,
                             int64_t *needed);

int rnng_graph_destroy(rnng_graph_t g);

/* One-shot host-to-host form with the count-then-fill contract: at most cap
 * edges are written to (ij, w), sorted by (i, j); *needed receives the total.
 * Returns RNNG_ERR_CAPACITY if cap < *needed (nothing partial is promised). */
int rnng_l2_eps_pairs_f32(const float *points, int64_t n, int32_t d, float eps, int32_t mode,
                          int64_t *ij, float *w, int64_t cap, int64_t *needed,
                          rnng_stats *stats);

int rnng_l2_eps_bipartite_f32(const float *a, int64_t na, const float *b, int64_t nb,
                              int32_t d, float eps, int32_t mode, int64_t *ij, float *w,
                              int64_t cap, int64_t *needed, rnng_stats *stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rnng_l2_eps_pairs_f32's self join, computed by its PLAIN
 * path's kernels, on device-resident buffers and with no allocation.
 * dev_points holds n x d floats (row-major) in device memory of the library's
 * device; the first *needed edges go to dev_ij[2e] = i, dev_ij[2e+1] = j and
 * dev_w[e] (device memory; either may be NULL), sorted by (i, j).  Every
 * intermediate lives in a fixed array of RNNG_NOALLOC_SCRATCH_BYTES in device
 * memory that each GPU holds until the process ends; one call handles at most
 * RNNG_NOALLOC_MAX_EDGES edges: a larger count returns RNNG_ERR_UNSUPPORTED
 * with *needed set, cap < *needed returns RNNG_ERR_CAPACITY (nothing written).
 * Every call on a GPU shares that array; the library serialises its calls, and
 * a call returns when its outputs are complete.  Pointers that are not device
 * memory of the library's device are RNNG_ERR_ARG.  stats (may be NULL):
 * pair_evals and edges. */
#define RNNG_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define RNNG_NOALLOC_MAX_EDGES (1ll << 21)
int rnng_l2_eps_pairs_f32_noalloc(const float *dev_points, int64_t n, int32_t d, float eps,
                                  int64_t *dev_ij, float *dev_w, int64_t cap, int64_t *needed,
                                  rnng_stats *stats);

/* Initial sizes (in edges and in filter candidates) of the buffers of the
 * next builds; 0 restores the library default.  Undersized buffers are safe:
 * the kernels keep counting past the end, and the call grows the buffer to
 * the exact count and runs again (count-then-fill), reported in
 * rnng_stats.refills.  The learned cost model uses this to size buffers. */
int rnng_set_initial_capacity(int64_t edges, int64_t candidates);

/* ------------------------------------------------------------------------ */
/* Edit-distance epsilon graph                                               */
/* ------------------------------------------------------------------------ */

typedef struct rnng_strings_s *rnng_strings_t;   /* device-resident string set */

/* Upload n strings given as a CSR: string i is atoms[offsets[i] ..
 * offsets[i+1]) (the layout PointContainer<char> keeps after read_seqs,
 * include/point.hpp:170-205).  Any byte values are allowed. */
int rnng_strings_create(const char *atoms, const int64_t *offsets, int64_t n,
                        rnng_strings_t *out);

/* A .seqs file straight into HBM: one string per line, a trailing '\r'
 * dropped, empty lines skipped, ids in file order (include/point.hpp:170-205);
 * strings [start, start + count) of the file (count < 0: to the end).  The
 * file travels in chunks through two pinned host buffers (the next chunk read
 * while the previous one is in flight); line splitting and the set's device
 * layout are built on the GPU.  RNNG_ERR_ARG for a range beyond the file. */
int rnng_strings_read_file(const char *path, int64_t start, int64_t count, rnng_strings_t *out);

/* Strings and atoms of a set; its atoms and offsets (n + 1, from 0) copied to
 * the host in id order (either may be NULL). */
int rnng_strings_size(rnng_strings_t s, int64_t *n, int64_t *atoms);
int rnng_strings_copy(rnng_strings_t s, char *atoms, int64_t *offsets);

int rnng_strings_destroy(rnng_strings_t s);

/* All pairs i < j whose Levenshtein distance (unit costs, the reference's
 * EditDistance functor, drivers/systolic.cpp:271-299) is <= eps; the weight
 * is the distance as a float.  Any eps and any string lengths. */
int rnng_edit_graph_build(rnng_strings_t s, float eps, rnng_graph_t *out, rnng_stats *stats);

/* All pairs (i, j), i in a, j in b, whose Levenshtein distance is <= eps, the
 * weight the distance as a float: the edit-distance form of
 * rnng_graph_build_bipartite (include/search.hpp:266-340 and :342-381 over
 * two string sets: cell x ghosts, a single query).  a and b may be the same
 * set (then every ordered pair, i == j included).  Limits as
 * rnng_edit_graph_build. */
int rnng_edit_graph_build_bipartite(rnng_strings_t a, rnng_strings_t b, float eps, rnng_graph_t *out,
                                    rnng_stats *stats);

/* One-shot host form with the count-then-fill contract of rnng_l2_eps_pairs_f32. */
int rnng_edit_eps_pairs(const char *atoms, const int64_t *offsets, int64_t n, float eps,
                        int64_t *ij, float *w, int64_t cap, int64_t *needed, rnng_stats *stats);

/* ------------------------------------------------------------------------ */
/* Vietoris-Rips filtration                                                  */
/* ------------------------------------------------------------------------ */

typedef struct rnng_filtration_s *rnng_filtration_t;   /* device-resident simplices */

/* Rips filtration of the L2 epsilon graph of ds up to dimension maxdim
 * (0..3): every vertex (value 0), edge (value = float L2 distance), triangle
 * and tetrahedron (value = max of its edges), with the reference's simplex ids
 * (include/simplex.hpp:5-19), sorted by (value, id) and unique by id -- the
 * merged output of the rips driver (include/point.hpp:1134-1267,
 * include/simplex.hpp:130-196).  Triangles need n <= 1,800,000 (60-bit ids);
 * tetrahedra n <= 55,111 (beyond it the reference's int64 binomial
 * v(v-1)(v-2)(v-3) overflows, include/utils.hpp:126-127); larger inputs give
 * RNNG_ERR_UNSUPPORTED.  The graph runs in RNNG_MODE_GRID for d <= 3,
 * RNNG_MODE_MFMA for d >= 16 and RNNG_MODE_PLAIN otherwise.  stats: pair
 * counters of the graph; filter_ms = graph, recheck_ms = CSR and cliques,
 * prep_ms = (value, id) ordering, total_ms = all. */
int rnng_rips_build(rnng_dataset_t ds, float eps, int32_t maxdim, rnng_filtration_t *out,
                    rnng_stats *stats);

/* Total number of simplices, and per dimension (per_dim[0..3], may be NULL). */
int rnng_filtration_size(rnng_filtration_t f, int64_t *n, int64_t *per_dim);

/* Copy the first cap simplices in filtration order (either array may be NULL). */
int rnng_filtration_copy(rnng_filtration_t f, float *value, int64_t *id, int64_t cap);

/* Relabels the vertices of a filtration built on n points (a Voronoi complex's
 * cell and ghost points; n must be that number): vertex v becomes ids[v]
 * (distinct, >= 0), every
 * simplex id is rebuilt from the relabelled vertices and the filtration is
 * re-sorted by (value, id) -- the global ids VoronoiComplex::build_filtration
 * gives its simplices (include/point.hpp:1218-1231).  vertex_flags (n bytes,
 * may be NULL): simplex_flags (may be NULL; one byte per simplex, in the new
 * order) receives their OR over each simplex's vertices (the interior flag,
 * :1163-1175).  Labels past the id limits of rnng_rips_build give
 * RNNG_ERR_UNSUPPORTED. */
int rnng_filtration_relabel(rnng_filtration_t f, const int64_t *ids, int64_t n, const uint8_t *vertex_flags,
                            uint8_t *simplex_flags);

int rnng_filtration_destroy(rnng_filtration_t f);

/* ------------------------------------------------------------------------ */
/* Voronoi landmark cells and ghost points                                   */
/* ------------------------------------------------------------------------ */

typedef struct rnng_voronoi_s *rnng_voronoi_t;   /* cells of a dataset (borrowed) */

/* Knuth's selection sampling of `size` ids from [0, range), ascending, with
 * the reference's engine and distribution (default_random_engine(seed),
 * uniform_real_distribution<double>(0, nextafter(1, max))); seed < 0 draws
 * one from std::random_device.  Host only.
 * Replaces: selection_sample, include/utils.hpp:1-29. */
int rnng_selection_sample(int64_t range, int64_t size, int64_t *sample, int32_t seed);

/* Nearest center of every point, reference-order FP32 distances, ties to the
 * later center; cell[i] in [0, m), dist[i] = that distance.  Host buffers.
 * Replaces: VoronoiDiagram ctor, include/point.hpp:535-559. */
int rnng_voronoi_assign_f32(const float *points, int64_t n, const float *centers, int64_t m,
                            int32_t d, int64_t *cell, float *dist);

/* Voronoi cells of ds around the points centers[0..m) (cell c = centers[c]):
 * assignment as rnng_voronoi_assign_f32, then the points grouped by cell in
 * ascending id with the center moved to the front.  ds must outlive the
 * handle.
 * Replaces: VoronoiDiagram ctor + coalesce_cells, include/point.hpp:535-758
 * (one GPU holds every cell, so there is no cell exchange). */
int rnng_voronoi_create(rnng_dataset_t ds, const int64_t *centers, int64_t m, rnng_voronoi_t *out);

/* As rnng_voronoi_create with m farthest-point landmarks: center 0 is point
 * `first`, center k+1 the point farthest (exact FP32 distance, ties to the
 * smaller id) from centers 0..k.  The landmark driver samples centers with
 * selection_sample instead (drivers/landmark.cpp:111-124); the edge set does
 * not depend on the centers. */
int rnng_voronoi_create_fps(rnng_dataset_t ds, int64_t m, int64_t first, rnng_voronoi_t *out);

/* Center point ids in cell order (m). */
int rnng_voronoi_centers(rnng_voronoi_t v, int64_t *ids);

/* Cell and distance to the own center of every point (either may be NULL). */
int rnng_voronoi_assignment(rnng_voronoi_t v, int64_t *cell, float *dist);

/* Cell CSR: cell_ptr[m+1], cell_ids[n] (coalesce order; either may be NULL).
 * Replaces: VoronoiCell membership after coalesce_cells. */
int rnng_voronoi_cells(rnng_voronoi_t v, int64_t *cell_ptr, int64_t *cell_ids);

/* Epsilon graph through the cells: ghost points (q joins cell b != cell(q)
 * when L2(center_b, q) <= fl(dist_q + 2 eps) and some member of b is within
 * eps of q), then every cell's member x member and member x ghost pairs
 * within eps, deduplicated to undirected edges i < j.  The edge set equals
 * the all-pairs graph's.  mode must be RNNG_MODE_MFMA (d <= 144).
 * Replaces: add_ghost_points_systolic (include/point.hpp:760-929), the
 * per-cell query loop (drivers/landmark.cpp:203-210) and the Graph ctor's
 * sort/unique (include/graph.hpp:59-116). */
int rnng_voronoi_graph_build(rnng_voronoi_t v, float eps, int32_t mode, rnng_graph_t *out,
                             rnng_stats *stats);

/* Per-handle switches of rnng_voronoi_graph_build (1 on, 0 off, -1 the
 * default: on unless RNNG_CELL_GRAPH / RNNG_CELL_TUNE is 0): use_graph replays
 * the cost-model-sized fill pass (filter, recheck, reference semantics,
 * per-cell counts) as a captured HIP graph; autotune times every tuned filter
 * configuration on every 8th work item at the next build and keeps the
 * fastest. */
int rnng_voronoi_set_options(rnng_voronoi_t v, int32_t use_graph, int32_t autotune);

/* Filter configuration in use (the BF16 variant of rnng_set_filter_kernels:
 * the autotuner times 17 = twelve waves sharing two LDS stages of 8 column
 * tiles, 13 / 14 = four waves with two stages of 2 / 3 tiles, 12 = four waves
 * with three stages of 2 tiles, or, when the kernels in use are the 16x16
 * forms and the k-step count is even, 19 and 18 (the 16x16 forms of 13 and
 * 17); xcd_order: the XCD-ordered work list) and the autotuner's timings in ms
 * of those variants on runs in landmark order, each without and with the XCD
 * order (8 values: 4 variants, or 19 and 18 in the first 4 and zeros after;
 * config_ms may be NULL; zeros if not tuned, and on integer data, whose int8
 * filter has one cell shape).
 * rnng_voronoi_schedule reports every configuration the tuner timed. */
int rnng_voronoi_tuning(rnng_voronoi_t v, int32_t *variant, int32_t *xcd_order, double *config_ms);

/* Schedule of the cell filter (no counterpart in the reference): the row
 * order of the layout (0 cells in landmark order, 1 cells along the centers'
 * nearest-neighbour chain, so that neighbouring cells share rows of the
 * layout; 2 flat: the points in their own order and every pair in the work,
 * for diagrams whose adjacency prunes little, runs only) and the work form
 * (0 runs of consecutive column tiles, one workgroup each; 1 each row block's
 * needed column tiles as one list, cut into items of up to 256 tiles; 2 those
 * items on a persistent grid of resident workgroups that take them from a
 * queue, heaviest first by tiles and the cost model's predicted candidates).  By default the autotuner times every (order, form, filter
 * shape) on every 8th item of the first build and keeps the fastest;
 * set_schedule forces an order and/or form (-1 leaves it to the tuner) from
 * the next build on.  rnng_voronoi_schedule reports the order and form in use
 * and the tuner's configurations (filter shape << 16 | order << 8 | form << 4 |
 * XCD order; the shape is the BF16 variant, or for integer data the int8 code,
 * 0 for the cell default) with
 * their times in ms (up to cap; *n_configs = how many were timed).  Every
 * schedule gives the same edges. */
int rnng_voronoi_set_schedule(rnng_voronoi_t v, int32_t order, int32_t sched);
int rnng_voronoi_schedule(rnng_voronoi_t v, int32_t *order, int32_t *sched, int32_t *configs, double *config_ms,
                          int32_t cap, int32_t *n_configs);

/* Ghosts found by the last rnng_voronoi_graph_build: their count, and the
 * (cell, point) pairs sorted by cell then point into pairs[2*cap].
 * Replaces: VoronoiCell::ghosts(), include/point.h:120. */
int rnng_voronoi_num_ghosts(rnng_voronoi_t v, int64_t *n);
int rnng_voronoi_ghosts(rnng_voronoi_t v, int64_t *pairs, int64_t cap);

int rnng_voronoi_destroy(rnng_voronoi_t v);

/* ------------------------------------------------------------------------ */
/* Cost model of the cell query (no counterpart in the reference)            */
/* ------------------------------------------------------------------------ */

/* A 4 -> 64 -> 64 -> 2 perceptron (ReLU, FP32) mapping per-cell features to
 * log-scaled per-cell candidate and edge counts.  Parameters are packed
 * W1[64][4], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] (4610 floats).  All
 * buffers are host memory; the work runs in the library's kernels. */
#define RNNG_MLP_NPARAMS 4610

/* y[batch][2] = the network on x[batch][4] (the inference kernel). */
int rnng_mlp_infer(const float *params, const float *x, int64_t batch, float *y);

/* Loss = mean over batch and outputs of w (y - target)^2, w = the
 * under-weight where y < target (an estimate below the count costs a second
 * filter pass, one above it only memory) and 1 elsewhere, and its gradient
 * with respect to the parameters (forward + backward kernels). */
int rnng_mlp_loss_grad(const float *params, const float *x, const float *target, int64_t batch,
                       float *loss, float *grad);

/* The under-weight of the loss (process-wide, for rnng_mlp_loss_grad and
 * rnng_mlp_train; default 4; 0 < w < 1e6). */
int rnng_mlp_set_under_weight(float w);
float rnng_mlp_under_weight(void);

/* `iters` full-batch Adam steps (lr, betas 0.9/0.999) on the GPU, updating
 * params in place; final_loss (may be NULL) = loss of the result. */
int rnng_mlp_train(float *params, const float *x, const float *target, int64_t batch, int32_t iters,
                   float lr, float *final_loss);

/* Cost model used by rnng_voronoi_graph_build to size its candidate buffer
 * (1.5 x the summed per-cell estimates: one filter pass; a second only if the
 * estimate was short).  NULL
 * removes it: builds then count first and fill in a second pass. */
int rnng_set_cost_model(const float *params);

/* Per-cell features (m x 4) and actual (candidates, edges) counts (m x 2) of
 * the last rnng_voronoi_graph_build, counted by the row point's cell: the
 * cost model's training data.  Features: log2(1+n)/16, log2(1+paired column
 * points)/24, log2(1+R/eps)/8 (R = largest member distance to the center),
 * log2(1+n*k)/24 (k = mean number of members within eps of the first
 * min(32, n) members); targets are log2(1+count)/24. */
int rnng_voronoi_cell_stats(rnng_voronoi_t v, float *features, float *counts);

/* Largest float T with sqrtf(T) <= eps (-1 when no value qualifies).  The
 * kernels compare squared distances against T; sqrtf is monotone, so
 * sqrtf(v) <= eps  <=>  v <= T. */
float rnng_sq_threshold(float eps);

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------ */

/* Runs one 32x32x16 BF16 MFMA with inputs that encode row and column
 * indices and reports, for each (lane, accumulator register), the output
 * row and column it holds: rows[lane*16 + r], cols[lane*16 + r]. */
int rnng_debug_mfma_layout(int32_t *rows, int32_t *cols);

/* MFMA mode on integer-valued data whose dimensions each span at most 255
 * uses the exact int8 filter unless disabled here (enable = 0: always BF16).
 * Same edges either way; a switch for comparisons. */
int rnng_set_int8_filter(int32_t enable);

/* Filter kernel shapes, process-wide (no counterpart in the reference: for
 * measurement and for tests that run every shape).  int8_code: 0 column tiles
 * staged through registers, 1/2 direct-to-LDS ring with three stages of 1/2
 * tiles, 3/4 two stages of 3/4 tiles, 5/6/7 two stages of 4/6/8 tiles shared
 * by eight waves of two row tiles, 8 two stages of 8 tiles shared by four waves
 * of four row tiles (two waves per SIMD), 9/10 the 16x16 MFMA forms
 * (v_mfma_i32_16x16x64_i8, 16-row fragment layout) of 7/3, 11/12 16x16 forms
 * with four row tiles per wave (four waves, two per SIMD) / three (twelve
 * waves, three per SIMD) (all-pairs builds; the cell path runs code 3 for
 * codes 5-8, and code 10 for 9-12).
 * bf16_variant 0..19 (11/12 three-stage ring, 13/14 two-stage ring of 2/3
 * tiles, 15 two stages of 4 tiles shared by 8 waves, 16/17 two stages of 4/8
 * tiles shared by 12 waves, 18/19 the 16x16 forms (v_mfma_f32_16x16x32_bf16)
 * of 17/13, the others register-staged or group-load shapes).  The 16x16 forms
 * need an even k-step count (int8: ceil(d / 32), BF16: ceil(d / 16)) and run
 * as their 32x32 forms otherwise.  A negative value leaves that setting
 * unchanged; the defaults (9, 19) come from RNNG_L2_I8_GROUP / RNNG_L2_VARIANT
 * when set.  Every shape gives the same edges. */
int rnng_set_filter_kernels(int32_t int8_code, int32_t bf16_variant);
int rnng_get_filter_kernels(int32_t *int8_code, int32_t *bf16_variant);

/* The XCD ordering of a filter work list, on the host (no device work):
 * items holds n work items of 4 int32 (row block, first column tile, end
 * column tile, flags); out receives *n_out = 8 ceil(n / 8) items in launch
 * order, workgroup b taking item (b mod 8) ceil(n / 8) + b / 8, so that XCD x
 * (which the hardware gives every 8th workgroup) runs the x-th contiguous
 * eighth of the list; positions past the list hold empty items (first tile
 * not below the end tile) that the filter kernels skip. */
int rnng_debug_xcd_permute(const int32_t *items, int64_t n, int32_t *out, int64_t cap, int64_t *n_out);

/* The cell filter's work lists by the host builder (rnng_debug_cell_work, no
 * device work) and by the device builder the graph build uses
 * (rnng_debug_cell_work_device): for m cells of cnt[c] points with the
 * m x ceil(m / 32) adjacency bitset adj (bit b of row a: cells a and b may hold
 * an edge), laid out in position order perm (position -> cell), with bt row
 * tiles per row block: start[c] = first slot of cell c (may be NULL); sched 0:
 * items (row block, first tile, end tile, 0) of runs of consecutive needed
 * cells' column tiles, cut at multiples of 256 tiles, in order of that chunk
 * then row block; 1 / 2: items (row block, first, end, 0) indexing ctl, each
 * row block's needed column tiles in ascending order, cut into items of at
 * most 256, in row-block order (2: by cost, heaviest first: tiles + the row
 * block's rb_cand (NULL: zeros) times the item's share of its tiles / 64).
 * *covered = pairs the tiles cover.  Both builders give the same arrays. */
int rnng_debug_cell_work(const int32_t *cnt, int32_t m, const uint32_t *adj, const int32_t *perm, int32_t bt,
                         int32_t sched, const float *rb_cand, int32_t *start, int32_t *items, int64_t cap_items,
                         int32_t *ctl, int64_t cap_ctl, int64_t *n_items, int64_t *n_ctl, int64_t *covered);
int rnng_debug_cell_work_device(const int32_t *cnt, int32_t m, const uint32_t *adj, const int32_t *perm,
                                int32_t bt, int32_t sched, const float *rb_cand, int32_t *start, int32_t *items,
                                int64_t cap_items, int32_t *ctl, int64_t cap_ctl, int64_t *n_items, int64_t *n_ctl,
                                int64_t *covered);

/* Runs one 32x32x32 int8 MFMA (the edit filter's) on caller data: a and b
 * hold 16 int8 per lane (64 lanes), d gets the 16 int32 accumulators of
 * each lane (started at zero): d[lane*16 + r]. */
int rnng_debug_mfma_i8(const int8_t *a, const int8_t *b, int32_t *d);

/* out[i] = the device square root the kernels use for edge weights, for
 * checking it against the host's correctly rounded sqrtf. */
int rnng_debug_sqrt(const float *in, float *out, int64_t n);

/* Copies nbytes (a multiple of 16) between two device buffers twice with a
 * plain vector-copy kernel and reports the second copy's time: a kernel of
 * exactly known HBM traffic for calibrating byte counters. */
int rnng_debug_copy(int64_t nbytes, double *ms);

#ifdef __cplusplus
}
#endif

#endif
