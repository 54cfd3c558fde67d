// Epsilon-range queries behind the reference's search API (include/search.h):
// CoverTree's build / radius_query / radius_query_batched /
// has_radius_neighbor and BruteForce::radius_query keep their signatures and
// their functor contract (called once per emitted pair, before the call
// returns) but run as GPU epsilon joins: no tree is built.  L2Distance<float>
// runs the BF16 MFMA filter with the exact FP32 recheck, EditDistance<char>
// the exact integer edit-distance kernels.
#ifndef RNNG_HOST_SEARCH_H_
#define RNNG_HOST_SEARCH_H_

#include "point.h"

namespace rnng_host {

struct GraphOwner {
    rnng_graph_t h = nullptr;
    ~GraphOwner() {
        if (h) rnng_graph_destroy(h);
    }
};

// (i, j, w) of a device graph, sorted by (i, j).
inline void graph_edges(rnng_graph_t g, std::vector<int64_t>& ij, std::vector<float>& w) {
    int64_t ne = 0;
    check(rnng_graph_num_edges(g, &ne), "graph_num_edges");
    ij.resize((size_t)(2 * ne));
    w.resize((size_t)ne);
    check(rnng_graph_copy(g, ij.data(), w.data(), ne, 1), "graph_copy");
}

// Pairs (i in A, j in B) within radius; self = (A is B): pairs i < j only.
template <class Atom>
void eps_join(const PointContainer<Atom>& a, const PointContainer<Atom>& b, bool self, Real radius,
              std::vector<int64_t>& ij, std::vector<float>& w) {
    GraphOwner g;
    if constexpr (std::is_same<Atom, float>::value) {
        // d <= 3: uniform grid (same edges, pairs of adjacent cells only)
        const int32_t mode = a.num_points() > 0 && a[0].size() <= 3 ? RNNG_MODE_GRID : RNNG_MODE_MFMA;
        if (self) check(rnng_graph_build(a.device_points(), radius, mode, &g.h, nullptr), "graph_build");
        else
            check(rnng_graph_build_bipartite(a.device_points(), b.device_points(), radius, RNNG_MODE_MFMA, &g.h,
                                             nullptr),
                  "graph_build_bipartite");
        graph_edges(g.h, ij, w);
    } else {
        if (self) {
            check(rnng_edit_graph_build(a.device_strings(), radius, &g.h, nullptr), "edit_graph_build");
            graph_edges(g.h, ij, w);
            return;
        }
        // cross pairs of the self join over A followed by B
        std::vector<char> atoms(a.atoms());
        atoms.insert(atoms.end(), b.atoms().begin(), b.atoms().end());
        IndexVector offs(a.atom_offsets());
        for (size_t k = 1; k < b.atom_offsets().size(); ++k) offs.push_back(a.num_atoms() + b.atom_offsets()[k]);
        const int64_t na = a.num_points(), n = na + b.num_points();
        StringsHandle s;
        check(rnng_strings_create(atoms.data(), offs.data(), n, &s.h), "strings_create");
        check(rnng_edit_graph_build(s.h, radius, &g.h, nullptr), "edit_graph_build");
        std::vector<int64_t> all;
        std::vector<float> aw;
        graph_edges(g.h, all, aw);
        ij.clear();
        w.clear();
        for (size_t e = 0; e < aw.size(); ++e) {
            int64_t i = all[2 * e], j = all[2 * e + 1];
            if (i > j) std::swap(i, j);
            if (i < na && j >= na) {
                ij.push_back(i);
                ij.push_back(j - na);
                w.push_back(aw[e]);
            }
        }
    }
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
    CoverTree() : cover(1.5), leaf_size(10) {}
    CoverTree(Real cover, Index leaf_size) : cover(cover), leaf_size(leaf_size) {}

    // Keeps the point set; the queries run against it on the GPU.
    template <class Atom, class Distance>
    void build(const PointContainer<Atom>&, Distance&) {
        rnng_host::require_gpu_metric<Distance, Atom>();
    }

    // functor(points[i], query, d) for every point within radius of query
    // (include/search.hpp:223-264).
    template <class Atom, class Distance, class Functor>
    void radius_query(const PointContainer<Atom>& points, Distance& distance, const Point<Atom>& query, Real radius,
                      const Functor& functor) const {
        PointContainer<Atom> batch;
        batch.push_back(query);
        radius_query_batched(points, distance, batch, radius,
                             [&](const Point<Atom>& p, const Point<Atom>&, Real d) { functor(p, query, d); });
    }

    // functor(points[i], batch[j], d) for every pair within radius
    // (include/search.hpp:266-340); when batch is the point set itself this
    // includes both orders of every pair and each point with itself.
    template <class Atom, class Distance, class Functor>
    void radius_query_batched(const PointContainer<Atom>& points, Distance&, const PointContainer<Atom>& batch,
                              Real radius, const Functor& functor) const {
        rnng_host::require_gpu_metric<Distance, Atom>();
        const bool self = &points == &batch;
        std::vector<int64_t> ij;
        std::vector<float> w;
        rnng_host::eps_join(points, batch, self, radius, ij, w);
        for (size_t e = 0; e < w.size(); ++e) {
            functor(points[ij[2 * e]], batch[ij[2 * e + 1]], w[e]);
            if (self) functor(points[ij[2 * e + 1]], batch[ij[2 * e]], w[e]);
        }
        if (self && radius >= 0)
            for (Index i = 0; i < points.num_points(); ++i) functor(points[i], batch[i], Real(0));
    }

    // Whether some point lies within radius of query (include/search.hpp:342-381).
    template <class Atom, class Distance>
    bool has_radius_neighbor(const PointContainer<Atom>& points, Distance& distance, const Point<Atom>& query,
                             Real radius) const {
        bool found = false;
        radius_query(points, distance, query, radius, [&](const Point<Atom>&, const Point<Atom>&, Real) { found = true; });
        return found;
    }

private:
    Real cover;
    Index leaf_size;
};

// Scan of all points (include/search.hpp:383-407), the semantic definition
// the CoverTree queries reproduce.
class BruteForce {
public:
    template <class Atom, class Distance>
    void build(const PointContainer<Atom>&, Distance&) {
        rnng_host::require_gpu_metric<Distance, Atom>();
    }
    template <class Atom, class Distance, class Functor>
    void radius_query(const PointContainer<Atom>& points, Distance& distance, const Point<Atom>& query, Real radius,
                      const Functor& functor) const {
        CoverTree().radius_query(points, distance, query, radius, functor);
    }
};

#endif
