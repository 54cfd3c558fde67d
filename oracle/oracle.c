/*
 * CPU oracle for the epsilon-neighbor-graph path.  TEST INFRASTRUCTURE ONLY:
 * the product never links or calls it.  Declarations, citations and the list
 * of checks that pin each function are in oracle.h.
 *
 * Build: see Makefile (gcc -O3 -ffp-contract=off, x86-64 baseline: no FMA).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------ */
/* L2 distance in the reference's order: drivers/systolic.cpp:253-269        */
/* ------------------------------------------------------------------------ */

float orc_l2sq(const float *p, const float *q, int64_t d)
{
    float val = 0.0f;
    for (int64_t k = 0; k < d; ++k) {
        float delta = p[k] - q[k];
        val = val + delta * delta;
    }
    return val;
}

float orc_l2(const float *p, const float *q, int64_t d)
{
    return sqrtf(orc_l2sq(p, q, d));
}

float orc_sq_threshold(float eps)
{
    if (!(eps >= 0.0f)) return -1.0f;          /* negative or NaN: no pair passes */
    if (isinf(eps)) return INFINITY;
    float v = (float)((double)eps * (double)eps);
    if (isinf(v)) v = FLT_MAX;
    while (v > 0.0f && sqrtf(v) > eps) v = nextafterf(v, -INFINITY);
    for (;;) {
        float up = nextafterf(v, INFINITY);
        if (isinf(up) || sqrtf(up) > eps) break;
        v = up;
    }
    return v;
}

/* ------------------------------------------------------------------------ */
/* Growable edge buffers used by the brute-force scans                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t *ij;     /* 2 per edge */
    float *w;
    int32_t *di;
    int64_t n, cap;
} edgebuf;

static void eb_push(edgebuf *b, int64_t i, int64_t j, float w, int32_t di)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 256;
        b->ij = (int64_t *)realloc(b->ij, sizeof(int64_t) * 2 * b->cap);
        b->w = (float *)realloc(b->w, sizeof(float) * b->cap);
        b->di = (int32_t *)realloc(b->di, sizeof(int32_t) * b->cap);
    }
    b->ij[2 * b->n] = i;
    b->ij[2 * b->n + 1] = j;
    b->w[b->n] = w;
    b->di[b->n] = di;
    b->n++;
}

static void eb_free(edgebuf *b)
{
    free(b->ij);
    free(b->w);
    free(b->di);
}

/* Concatenate chunk buffers in chunk order into the caller's arrays. */
static int64_t eb_gather(edgebuf *chunks, int64_t nchunks, int64_t *ij, float *w,
                         int32_t *di, int64_t cap)
{
    int64_t total = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
        edgebuf *b = &chunks[c];
        for (int64_t e = 0; e < b->n; ++e, ++total) {
            if (total < cap) {
                if (ij) { ij[2 * total] = b->ij[2 * e]; ij[2 * total + 1] = b->ij[2 * e + 1]; }
                if (w) w[total] = b->w[e];
                if (di) di[total] = b->di[e];
            }
        }
        eb_free(b);
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* Brute-force L2 epsilon range: include/search.hpp:389-407                 */
/* ------------------------------------------------------------------------ */

/* Eight pairs at a time.  Each vector lane carries one pair and accumulates
 * over k = 0..d-1 in sequence, so every lane's value is bit-identical to
 * orc_l2sq: lanes never mix, and -ffp-contract=off keeps sub, mul and add
 * separately rounded. */
typedef float v8f __attribute__((vector_size(32)));

#define ROW_CHUNK 16

int64_t orc_l2_eps_pairs(const float *X, int64_t nx, const float *Y, int64_t ny,
                         int64_t d, float eps, int self_join, int nthreads,
                         int64_t *ij, float *w, int64_t cap)
{
    if (nx <= 0 || ny <= 0) return 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();

    /* Y transposed in blocks of 8 columns: Yt[(jb*d + k)*8 + lane]. */
    int64_t nyb = (ny + 7) / 8;
    float *Yt = (float *)aligned_alloc(32, sizeof(float) * (size_t)(nyb * d * 8));
    for (int64_t jb = 0; jb < nyb; ++jb)
        for (int64_t k = 0; k < d; ++k)
            for (int l = 0; l < 8; ++l) {
                int64_t j = jb * 8 + l;
                Yt[(jb * d + k) * 8 + l] = (j < ny) ? Y[j * d + k] : 0.0f;
            }

    int64_t nchunks = (nx + ROW_CHUNK - 1) / ROW_CHUNK;
    edgebuf *chunks = (edgebuf *)calloc((size_t)nchunks, sizeof(edgebuf));

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t c = 0; c < nchunks; ++c) {
        int64_t i0 = c * ROW_CHUNK, i1 = i0 + ROW_CHUNK < nx ? i0 + ROW_CHUNK : nx;
        for (int64_t i = i0; i < i1; ++i) {
            const float *x = X + i * d;
            int64_t jstart = self_join ? i + 1 : 0;
            int64_t jb = jstart / 8;
            for (; jb < nyb; jb += 2) {
                v8f a0 = {0}, a1 = {0};
                const float *y0 = Yt + jb * d * 8;
                const float *y1 = (jb + 1 < nyb) ? Yt + (jb + 1) * d * 8 : y0;
                for (int64_t k = 0; k < d; ++k) {
                    float xk = x[k];
                    v8f xv = {xk, xk, xk, xk, xk, xk, xk, xk};
                    v8f t0 = xv - *(const v8f *)(y0 + k * 8);
                    v8f t1 = xv - *(const v8f *)(y1 + k * 8);
                    a0 = a0 + t0 * t0;
                    a1 = a1 + t1 * t1;
                }
                for (int l = 0; l < 16; ++l) {
                    int64_t j = jb * 8 + l;
                    if (l >= 8 && jb + 1 >= nyb) break;
                    if (j < jstart || j >= ny) continue;
                    float v = (l < 8) ? a0[l] : a1[l - 8];
                    float dist = sqrtf(v);
                    if (dist <= eps) eb_push(&chunks[c], i, j, dist, 0);
                }
            }
        }
    }

    free(Yt);
    int64_t total = eb_gather(chunks, nchunks, ij, w, NULL, cap);
    free(chunks);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Edit distance: drivers/systolic.cpp:271-299                               */
/* ------------------------------------------------------------------------ */

int64_t orc_edit_distance(const char *s, int64_t m, const char *t, int64_t n)
{
    int64_t *v0 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t *v1 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i <= n; ++i) v0[i] = i;
    for (int64_t i = 0; i < m; ++i) {
        v1[0] = i + 1;
        for (int64_t j = 0; j < n; ++j) {
            int64_t del = v0[j + 1] + 1;
            int64_t ins = v1[j] + 1;
            int64_t sub = (s[i] == t[j]) ? v0[j] : v0[j] + 1;
            int64_t best = ins < sub ? ins : sub;
            v1[j + 1] = del < best ? del : best;
        }
        int64_t *tmp = v0; v0 = v1; v1 = tmp;
    }
    int64_t r = v0[n];
    free(v0);
    free(v1);
    return r;
}

int64_t orc_edit_eps_pairs(const char *atoms, const int64_t *offsets, int64_t n,
                           float eps, int nthreads, int64_t *ij, int32_t *dist,
                           int64_t cap)
{
    if (n <= 0) return 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int64_t nchunks = (n + ROW_CHUNK - 1) / ROW_CHUNK;
    edgebuf *chunks = (edgebuf *)calloc((size_t)nchunks, sizeof(edgebuf));

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t c = 0; c < nchunks; ++c) {
        int64_t i0 = c * ROW_CHUNK, i1 = i0 + ROW_CHUNK < n ? i0 + ROW_CHUNK : n;
        for (int64_t i = i0; i < i1; ++i) {
            int64_t mi = offsets[i + 1] - offsets[i];
            for (int64_t j = i + 1; j < n; ++j) {
                int64_t nj = offsets[j + 1] - offsets[j];
                int64_t lb = mi > nj ? mi - nj : nj - mi;
                if ((float)lb > eps) continue;   /* |m-n| <= distance: exact skip */
                int64_t e = orc_edit_distance(atoms + offsets[i], mi, atoms + offsets[j], nj);
                if ((float)e <= eps) eb_push(&chunks[c], i, j, (float)e, (int32_t)e);
            }
        }
    }
    int64_t total = eb_gather(chunks, nchunks, ij, NULL, dist, cap);
    free(chunks);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Voronoi assignment: include/point.hpp:535-559                             */
/* ------------------------------------------------------------------------ */

void orc_voronoi_assign(const float *pts, int64_t n, const float *centers, int64_t m,
                        int64_t d, int nthreads, int64_t *cell, float *dist)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        float best = FLT_MAX;
        int64_t bc = 0;
        for (int64_t c = 0; c < m; ++c) {
            float dd = orc_l2(centers + c * d, pts + i * d, d);
            if (dd <= best) { best = dd; bc = c; }
        }
        cell[i] = bc;
        dist[i] = best;
    }
}

/* ------------------------------------------------------------------------ */
/* Farthest-point landmarks (no reference counterpart: the landmark driver    */
/* samples centers, drivers/landmark.cpp:111-124).  Center 0 = first; center  */
/* k+1 = argmax of the distance to the nearest chosen center, ties to the     */
/* smaller index.  The running (dist, cell) update uses the `<=` rule of      */
/* include/point.hpp:550-556, so after the last center (cell, dist) is the    */
/* Voronoi assignment of the centers in selection order.                      */
/* ------------------------------------------------------------------------ */

void orc_fps(const float *pts, int64_t n, int64_t d, int64_t m, int64_t first, int nthreads,
             int64_t *centers, int64_t *cell, float *dist)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    centers[0] = first;
    for (int64_t k = 0; k < m; ++k) {
        const float *c = pts + centers[k] * d;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t i = 0; i < n; ++i) {
            float dd = orc_l2(c, pts + i * d, d);
            float md = k == 0 ? FLT_MAX : dist[i];
            if (dd <= md) { md = dd; cell[i] = k; }
            dist[i] = md;
        }
        if (k + 1 < m) {
            int64_t bi = 0;
            for (int64_t i = 1; i < n; ++i)
                if (dist[i] > dist[bi]) bi = i;
            centers[k + 1] = bi;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Ghost criterion: include/point.hpp:894-918                                */
/* ------------------------------------------------------------------------ */

int64_t orc_ghosts(const float *pts, int64_t n, int64_t d, const int64_t *cell,
                   const float *delta, const float *centers, int64_t m, float eps,
                   int nthreads, int64_t *out, int64_t cap)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    /* CSR of cell members, members in ascending point order. */
    int64_t *cnt = (int64_t *)calloc((size_t)(m + 1), sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) cnt[cell[i] + 1]++;
    for (int64_t c = 0; c < m; ++c) cnt[c + 1] += cnt[c];
    int64_t *members = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    for (int64_t c = 0; c < m; ++c) fill[c] = cnt[c];
    for (int64_t i = 0; i < n; ++i) members[fill[cell[i]]++] = i;
    free(fill);

    /* One chunk per cell keeps the output sorted by (cell, point). */
    edgebuf *chunks = (edgebuf *)calloc((size_t)(m ? m : 1), sizeof(edgebuf));
    const float two_eps = 2.0f * eps;

#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t c = 0; c < m; ++c) {
        if (cnt[c + 1] == cnt[c]) continue;
        const float *ctr = centers + c * d;
        for (int64_t q = 0; q < n; ++q) {
            if (cell[q] == c) continue;
            const float *qp = pts + q * d;
            float reach = delta[q] + two_eps;
            if (!(orc_l2(ctr, qp, d) <= reach)) continue;
            for (int64_t e = cnt[c]; e < cnt[c + 1]; ++e) {
                if (orc_l2(pts + members[e] * d, qp, d) <= eps) {
                    eb_push(&chunks[c], c, q, 0.0f, 0);
                    break;
                }
            }
        }
    }
    int64_t total = eb_gather(chunks, m, out, NULL, NULL, cap);
    free(chunks);
    free(members);
    free(cnt);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Simplex ids: include/utils.hpp:116-144, include/simplex.hpp:5-19          */
/* ------------------------------------------------------------------------ */

int64_t orc_binom(int64_t n, int64_t k)
{
    if (k > n) return 0;
    if (k < 0) return 0;
    __int128 r = 1;
    /* Multiplicative form: each partial product r * (n-k+i) / i is an integer. */
    for (int64_t i = 1; i <= k; ++i) r = r * (__int128)(n - k + i) / i;
    return (int64_t)r;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_simplex_id(const int64_t *verts, int64_t k)
{
    int64_t v[16];
    for (int64_t i = 0; i < k; ++i) v[i] = verts[i];
    qsort(v, (size_t)k, sizeof(int64_t), cmp_i64);
    uint64_t p = (uint64_t)(k - 1), uid = 0;
    for (int64_t i = k - 1; i >= 0; --i) uid += (uint64_t)orc_binom(v[i], i + 1);
    return (int64_t)(uid | (p << 60));
}

/* ------------------------------------------------------------------------ */
/* Rips filtration: include/point.hpp:1134-1267, include/simplex.hpp:130-196 */
/* ------------------------------------------------------------------------ */

typedef struct { float value; int64_t id; } wsimplex;

static int cmp_ws(const void *a, const void *b)
{
    const wsimplex *x = (const wsimplex *)a, *y = (const wsimplex *)b;
    if (x->value < y->value) return -1;
    if (x->value > y->value) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

typedef struct {
    const int64_t *rowptr, *col;   /* upper adjacency: neighbours v > u, ascending */
    const float *wt;
    int64_t maxsize;
    wsimplex *buf;
    int64_t nbuf, capbuf;
} clique_ctx;

static float edge_weight(const clique_ctx *cx, int64_t u, int64_t v)
{
    int64_t lo = cx->rowptr[u], hi = cx->rowptr[u + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (cx->col[mid] < v) lo = mid + 1; else hi = mid;
    }
    return cx->wt[lo];
}

static void push_ws(clique_ctx *cx, float value, int64_t id)
{
    if (cx->nbuf == cx->capbuf) {
        cx->capbuf = cx->capbuf ? 2 * cx->capbuf : 1024;
        cx->buf = (wsimplex *)realloc(cx->buf, sizeof(wsimplex) * (size_t)cx->capbuf);
    }
    cx->buf[cx->nbuf].value = value;
    cx->buf[cx->nbuf].id = id;
    cx->nbuf++;
}

/* Extend the clique cur[0..k) (ascending) by candidates cand[0..nc), which are
 * common upper neighbours of every vertex in cur, ascending. */
static void extend(clique_ctx *cx, int64_t *cur, int64_t k, float val,
                   const int64_t *cand, int64_t nc)
{
    int64_t *next = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nc ? nc : 1));
    for (int64_t a = 0; a < nc; ++a) {
        int64_t v = cand[a];
        /* value of the new simplex: max over its facets, which equals the max
         * over all its edges; float max is exact so the order is irrelevant */
        float nv = val;
        for (int64_t i = 0; i < k; ++i) {
            float we = edge_weight(cx, cur[i], v);
            if (we > nv) nv = we;
        }
        cur[k] = v;
        push_ws(cx, nv, orc_simplex_id(cur, k + 1));
        if (k + 1 < cx->maxsize) {
            int64_t nn = 0;
            for (int64_t b = a + 1; b < nc; ++b) {
                int64_t x = cand[b];
                /* x adjacent to v?  (v < x) */
                int64_t lo = cx->rowptr[v], hi = cx->rowptr[v + 1];
                while (lo < hi) {
                    int64_t mid = (lo + hi) / 2;
                    if (cx->col[mid] < x) lo = mid + 1; else hi = mid;
                }
                if (lo < cx->rowptr[v + 1] && cx->col[lo] == x) next[nn++] = x;
            }
            if (nn) extend(cx, cur, k + 1, nv, next, nn);
        }
    }
    free(next);
}

int64_t orc_rips_filtration(const float *pts, int64_t n, int64_t d, float eps,
                            int64_t maxdim, int nthreads, float *value, int64_t *id,
                            int64_t cap)
{
    if (n <= 0) return 0;
    int64_t ne = orc_l2_eps_pairs(pts, n, pts, n, d, eps, 1, nthreads, NULL, NULL, 0);
    int64_t *ij = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)(ne ? ne : 1));
    float *w = (float *)malloc(sizeof(float) * (size_t)(ne ? ne : 1));
    orc_l2_eps_pairs(pts, n, pts, n, d, eps, 1, nthreads, ij, w, ne);

    /* pairs come out row-major (u ascending, v ascending): that is the CSR */
    int64_t *rowptr = (int64_t *)calloc((size_t)(n + 1), sizeof(int64_t));
    int64_t *col = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ne ? ne : 1));
    for (int64_t e = 0; e < ne; ++e) { rowptr[ij[2 * e] + 1]++; col[e] = ij[2 * e + 1]; }
    for (int64_t u = 0; u < n; ++u) rowptr[u + 1] += rowptr[u];

    clique_ctx cx = { rowptr, col, w, maxdim + 1, NULL, 0, 0 };
    int64_t cur[16];
    for (int64_t u = 0; u < n; ++u) {
        cur[0] = u;
        push_ws(&cx, 0.0f, orc_simplex_id(cur, 1));
        if (maxdim >= 1)
            extend(&cx, cur, 1, 0.0f, col + rowptr[u], rowptr[u + 1] - rowptr[u]);
    }

    qsort(cx.buf, (size_t)cx.nbuf, sizeof(wsimplex), cmp_ws);
    int64_t out = 0;
    for (int64_t s = 0; s < cx.nbuf; ++s) {
        if (out > 0 && cx.buf[s].id == cx.buf[out - 1].id) continue;
        cx.buf[out++] = cx.buf[s];
    }
    for (int64_t s = 0; s < out && s < cap; ++s) {
        if (value) value[s] = cx.buf[s].value;
        if (id) id[s] = cx.buf[s].id;
    }
    free(cx.buf);
    free(rowptr);
    free(col);
    free(ij);
    free(w);
    return out;
}
