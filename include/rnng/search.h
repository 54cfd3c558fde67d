// Epsilon-range queries behind the reference's search API (include/search.h):
// CoverTree's build / radius_query / radius_query_batched /
// has_radius_neighbor and BruteForce::radius_query keep their signatures,
// return values (the number of pairs found) and functor contract (called once
// per found pair, before the call returns), but run as GPU epsilon joins: no
// tree is built.  L2Distance<float> runs the MFMA filter with its exact
// recheck (or the uniform grid for d <= 3), EditDistance<char> the exact
// edit-distance kernels.  A Voronoi cell's queries against itself and its
// ghosts at the radius its diagram was built for come from that one batched
// build.
#ifndef RNNG_HOST_SEARCH_H_
#define RNNG_HOST_SEARCH_H_

#include "point.h"

namespace rnng_host {

// One device buffer per thread for single float queries, overwritten by each
// query: no device allocation per query once it exists for the dimension.
// Never freed (it lives until the process ends, past the runtime's teardown).
struct QueryBuffer {
    rnng_dataset_t h = nullptr;
    int32_t d = 0;
    rnng_dataset_t set(const float* x, int32_t dim) {
        if (h && d == dim) {
            check(rnng_dataset_upload_f32(h, x, 1), "dataset_upload_f32");
            return h;
        }
        if (h) rnng_dataset_destroy(h);
        h = nullptr;
        check(rnng_dataset_create_f32(x, 1, dim, &h), "dataset_create_f32");
        d = dim;
        return h;
    }
};
inline QueryBuffer& query_buffer() {
    static thread_local QueryBuffer* q = new QueryBuffer();
    return *q;
}

// Pairs (i, j), i < j, of one container within radius.
template <class Atom>
void self_join(const PointContainer<Atom>& a, Real radius, std::vector<int64_t>& ij, std::vector<float>& w) {
    GraphOwner g;
    if constexpr (std::is_same<Atom, float>::value) {
        // d <= 3: uniform grid (same edges, pairs of adjacent cells only)
        const int32_t mode = a.num_points() > 0 && a.dim() <= 3 ? RNNG_MODE_GRID : RNNG_MODE_MFMA;
        check(rnng_graph_build(a.device_points(), radius, mode, &g.h, nullptr), "graph_build");
    } else {
        check(rnng_edit_graph_build(a.device_strings(), radius, &g.h, nullptr), "edit_graph_build");
    }
    graph_edges(g.h, ij, w);
}

// Pairs (i in a, j in b) within radius.
template <class Atom>
void cross_join(const PointContainer<Atom>& a, const PointContainer<Atom>& b, Real radius, std::vector<int64_t>& ij,
                std::vector<float>& w) {
    ij.clear();
    w.clear();
    if (a.num_points() == 0 || b.num_points() == 0) return;
    GraphOwner g;
    if constexpr (std::is_same<Atom, float>::value) {
        if (a.dim() != b.dim()) fail("query points of another dimension");
        check(rnng_graph_build_bipartite(a.device_points(), b.device_points(), radius, RNNG_MODE_MFMA, &g.h, nullptr),
              "graph_build_bipartite");
    } else {
        check(rnng_edit_graph_build_bipartite(a.device_strings(), b.device_strings(), radius, &g.h, nullptr),
              "edit_graph_build_bipartite");
    }
    graph_edges(g.h, ij, w);
}

// Points of a within radius of one query point: (i, 0) pairs.
template <class Atom>
void query_join(const PointContainer<Atom>& a, const Point<Atom>& q, Real radius, std::vector<int64_t>& ij,
                std::vector<float>& w) {
    ij.clear();
    w.clear();
    if (a.num_points() == 0) return;
    GraphOwner g;
    if constexpr (std::is_same<Atom, float>::value) {
        if (q.size() != a.dim()) fail("query point of another dimension");
        rnng_dataset_t qd = query_buffer().set(q.begin(), (int32_t)q.size());
        // one column: the plain kernel's single launch decides every pair
        check(rnng_graph_build_bipartite(a.device_points(), qd, radius, RNNG_MODE_PLAIN, &g.h, nullptr),
              "graph_build_bipartite");
    } else {
        const int64_t offs[2] = {0, q.size()};
        StringsHandle s;
        check(rnng_strings_create(q.begin(), offs, 1, &s.h), "strings_create");
        check(rnng_edit_graph_build_bipartite(a.device_strings(), s.h, radius, &g.h, nullptr),
              "edit_graph_build_bipartite");
    }
    graph_edges(g.h, ij, w);
}

template <class Distance, class Atom>
constexpr void require_gpu_metric() {
    static_assert((is_l2<Distance>() && std::is_same<Atom, float>::value) ||
                      (is_edit<Distance>() && std::is_same<Atom, char>::value),
                  "GPU epsilon queries exist for L2Distance<float> and EditDistance<char>");
}

}  // namespace rnng_host

class CoverTree {
public:
    CoverTree(Real cover = 1.5, Index leaf_size = 1) : cover(cover), leaf_size(leaf_size) {}

    // Keeps nothing: the queries run against the point set on the GPU.
    template <class Atom, class Distance>
    void build(const PointContainer<Atom>&, Distance&) {
        rnng_host::require_gpu_metric<Distance, Atom>();
    }

    // functor(points[i], query, d) for every point within radius of query;
    // returns their number (include/search.hpp:223-264).
    template <class Atom, class Distance, class Functor>
    Index radius_query(const PointContainer<Atom>& points, Distance&, const Point<Atom>& query, Real radius,
                       const Functor& functor) const {
        rnng_host::require_gpu_metric<Distance, Atom>();
        std::vector<int64_t> ij;
        std::vector<float> w;
        rnng_host::query_join(points, query, radius, ij, w);
        for (size_t e = 0; e < w.size(); ++e) functor(points[ij[2 * e]], query, w[e]);
        return (Index)w.size();
    }

    // functor(points[i], batch[j], d) for every pair within radius; returns
    // their number (include/search.hpp:266-340).  When batch is the point set
    // itself that includes both orders of every pair and each point with
    // itself at distance 0.
    template <class Atom, class Distance, class Functor>
    Index radius_query_batched(const PointContainer<Atom>& points, Distance&, const PointContainer<Atom>& batch,
                               Real radius, const Functor& functor) const {
        rnng_host::require_gpu_metric<Distance, Atom>();
        std::vector<int64_t> ij;
        std::vector<float> w;
        if (&points == &batch) {
            if (!(radius >= 0)) return 0;
            rnng_host::self_join(points, radius, ij, w);
            for (size_t e = 0; e < w.size(); ++e) {
                functor(points[ij[2 * e]], batch[ij[2 * e + 1]], w[e]);
                functor(points[ij[2 * e + 1]], batch[ij[2 * e]], w[e]);
            }
            for (Index i = 0; i < points.num_points(); ++i) functor(points[i], batch[i], Real(0));
            return (Index)(2 * w.size()) + points.num_points();
        }
        rnng_host::cross_join(points, batch, radius, ij, w);
        for (size_t e = 0; e < w.size(); ++e) functor(points[ij[2 * e]], batch[ij[2 * e + 1]], w[e]);
        return (Index)w.size();
    }

    // A Voronoi cell against itself or its ghosts at the radius its diagram's
    // batched build used: the same pairs, read from that build.
    template <class Atom, class Distance, class Functor>
    Index radius_query_batched(const VoronoiCell<Atom>& cell, Distance& distance, const PointContainer<Atom>& batch,
                               Real radius, const Functor& functor) const {
        const rnng_host::CellPairs* cp = cell.served_pairs();
        const PointContainer<Atom>& members = cell;
        if (!cp || cp->radius != radius || (&batch != &members && &batch != &cell.ghosts()))
            return radius_query_batched(members, distance, batch, radius, functor);
        Index found = 0;
        if (&batch == &members) {
            for (const auto& [i, j, d] : cp->members) {
                functor(cell[i], cell[j], d);
                functor(cell[j], cell[i], d);
            }
            for (Index i = 0; i < cell.num_points(); ++i) functor(cell[i], cell[i], Real(0));
            found = (Index)(2 * cp->members.size()) + cell.num_points();
        } else {
            for (const auto& [i, g, d] : cp->ghosts) functor(cell[i], batch[g], d);
            found = (Index)cp->ghosts.size();
        }
        return found;
    }

    // Whether some point lies within radius of query (include/search.hpp:342-381).
    template <class Atom, class Distance>
    bool has_radius_neighbor(const PointContainer<Atom>& points, Distance& distance, const Point<Atom>& query,
                             Real radius) const {
        return radius_query(points, distance, query, radius, [](const Point<Atom>&, const Point<Atom>&, Real) {}) > 0;
    }

private:
    Real cover;
    Index leaf_size;
};

// Scan of all points (include/search.hpp:383-407), the semantic definition
// the CoverTree queries reproduce.
class BruteForce {
public:
    BruteForce() {}
    template <class Atom, class Distance>
    void build(const PointContainer<Atom>&, Distance&) {
        rnng_host::require_gpu_metric<Distance, Atom>();
    }
    template <class Atom, class Distance, class Functor>
    Index radius_query(const PointContainer<Atom>& points, Distance& distance, const Point<Atom>& query, Real radius,
                       const Functor& functor) const {
        return CoverTree().radius_query(points, distance, query, radius, functor);
    }
};

#endif
