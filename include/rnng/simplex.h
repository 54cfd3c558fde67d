// Simplices and the merged filtration file (include/simplex.h, simplex.hpp).
#ifndef RNNG_HOST_SIMPLEX_H_
#define RNNG_HOST_SIMPLEX_H_

#include <sstream>

#include "utils.h"

namespace rnng_host {

// C(n, k) for the simplex ids: exact in 128 bits, as the reference's int64
// Binom for every id below 2^60 (include/utils.hpp:116-144).
inline Index binom(Index n, Index k) {
    if (k < 0 || n < k) return 0;
    if (k == 0 || n == k) return 1;
    __int128 r = 1;
    for (Index i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return (Index)r;
}

}  // namespace rnng_host

// A simplex's id: sum_i C(v_i, i + 1) over its sorted vertices, with the
// dimension in the top 4 bits (include/simplex.hpp:5-19).
struct Simplex {
    Simplex() : id(0) {}
    explicit Simplex(Index id) : id(id) {}
    explicit Simplex(const IndexVector& verts) {
        IndexVector v(verts);
        std::sort(v.begin(), v.end());
        const uint64_t p = v.size() - 1;
        uint64_t uid = 0;
        for (Index i = (Index)p; i >= 0; --i) uid += (uint64_t)rnng_host::binom(v[(size_t)i], i + 1);
        id = (Index)(uid | (p << 60));
    }
    Index getid() const { return id; }
    Index getuid() const { return (Index)((uint64_t)id & 0xFFFFFFFFFFFFFFFull); }
    Index getdim() const { return (Index)(((uint64_t)id >> 60) & 0xF); }
    // Vertices from the id (combinatorial number system), ascending.
    IndexVector getverts(Index n) const {
        uint64_t uid = (uint64_t)getuid();
        IndexVector verts;
        for (Index i = getdim(); i >= 0; --i) {
            Index lo = 0, hi = n;   // largest l < n with C(l, i + 1) <= uid
            while (lo < hi) {
                const Index mid = lo + ((hi - lo) >> 1);
                if ((uint64_t)rnng_host::binom(mid, i + 1) > uid) hi = mid;
                else lo = mid + 1;
            }
            const Index l = hi - 1;
            verts.push_back(l);
            uid -= (uint64_t)rnng_host::binom(l, i + 1);
        }
        std::sort(verts.begin(), verts.end());
        return verts;
    }
    // The facets: each vertex left out in turn (include/simplex.hpp:81-111).
    void get_facets(std::vector<Simplex>& facets, Index n) const {
        facets.clear();
        const IndexVector v = getverts(n);
        for (size_t i = 0; i < v.size() && v.size() > 1; ++i) {
            IndexVector f;
            for (size_t j = 0; j < v.size(); ++j)
                if (j != i) f.push_back(v[j]);
            facets.emplace_back(f);
        }
    }
    void get_facet_ids(IndexVector& ids, Index n) const {
        std::vector<Simplex> facets;
        get_facets(facets, n);
        ids.clear();
        for (const auto& f : facets) ids.push_back(f.getid());
    }
    std::string get_simplex_repr(Index n) const {
        const IndexVector v = getverts(n);
        std::stringstream ss;
        ss << "<";
        for (size_t i = 0; i + 1 < v.size(); ++i) ss << v[i] << ",";
        ss << v.back() << ">";
        return ss.str();
    }
    Index id;
};

struct WeightedSimplex : public Simplex {
    WeightedSimplex() : Simplex(), value(0), interior(true) {}
    WeightedSimplex(Index id, Real value) : Simplex(id), value(value), interior(true) {}
    WeightedSimplex(const IndexVector& verts, bool interior) : Simplex(verts), value(0), interior(interior) {}
    friend bool operator<(const WeightedSimplex& a, const WeightedSimplex& b) {
        return std::tie(a.value, a.id) < std::tie(b.value, b.id);
    }
    friend bool operator==(const WeightedSimplex& a, const WeightedSimplex& b) { return a.id == b.id; }
    friend bool operator!=(const WeightedSimplex& a, const WeightedSimplex& b) { return a.id != b.id; }
    Real getvalue() const { return value; }
    Real value;
    bool interior;
};

// Sort by (value, id), drop repeated ids, write "%f\t%lld" (use_ids) or
// "%f\t<v0,...>" lines (include/simplex.hpp:130-196, one rank).
inline void merge_and_write_filtration(const char* fname, const std::vector<WeightedSimplex>& simplices,
                                       Index num_vertices, bool use_ids) {
    std::vector<WeightedSimplex> s(simplices);
    if (!std::is_sorted(s.begin(), s.end())) std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    FILE* f = std::fopen(fname, "w");
    if (!f) {
        std::fprintf(stderr, "rnng: cannot open %s\n", fname);
        std::abort();
    }
    for (const auto& x : s) {
        if (use_ids) std::fprintf(f, "%f\t%lld\n", x.value, (long long)x.id);
        else std::fprintf(f, "%f\t%s\n", x.value, x.get_simplex_repr(num_vertices).c_str());
    }
    std::fclose(f);
}

// The communicator form (rank 0 gathers and writes): one process writes.
inline void merge_and_write_filtration(const char* fname, const std::vector<WeightedSimplex>& simplices,
                                       Index num_vertices, bool use_ids, SingleComm) {
    merge_and_write_filtration(fname, simplices, num_vertices, use_ids);
}

#endif
