// The reference drivers' bodies on the host layer (include/rnng/*.h), one
// process: the same classes, calls and call order as drivers/systolic.cpp
// (ring of single radius queries), drivers/landmark.cpp (per-cell batched
// queries against the cell and its ghosts) and drivers/rips.cpp (one complex
// per cell, merged filtration), with the communicator replaced by the
// one-process token.  tests/test_gpu_driver_bodies.py runs it on the GPU and
// compares the files it writes with the reference's own outputs.
//
//   driver_bodies systolic <points|seqs> <radius> <out> <l2|edit>
//   driver_bodies landmark <points> <radius> <centers> <seed> <out> <served|generic>
//   driver_bodies rips <points> <radius> <maxdim> <centers> <seed> <prefix> <ids|verts>
#include <cmath>
#include <cstdio>
#include <string>

#include "rnng/graph.h"
#include "rnng/search.h"

template <class Atom>
struct L2Distance {
    Index dist_comps = 0;
    Real operator()(const Point<Atom>& p, const Point<Atom>& q) {
        ++dist_comps;
        volatile Real acc = 0;
        for (Index i = 0; i < p.size(); ++i) {
            const Real t = p[i] - q[i];
            const Real t2 = t * t;
            acc = acc + t2;
        }
        return std::sqrt((Real)acc);
    }
};

template <class Atom>
struct EditDistance {
    Index dist_comps = 0;
    Real operator()(const Point<Atom>& s, const Point<Atom>& t) {
        ++dist_comps;
        const Index m = s.size(), n = t.size();
        std::vector<Index> prev((size_t)n + 1), cur((size_t)n + 1);
        for (Index j = 0; j <= n; ++j) prev[(size_t)j] = j;
        for (Index i = 1; i <= m; ++i) {
            cur[0] = i;
            for (Index j = 1; j <= n; ++j)
                cur[(size_t)j] = std::min(std::min(prev[(size_t)j] + 1, cur[(size_t)j - 1] + 1),
                                          prev[(size_t)j - 1] + (s[i - 1] == t[j - 1] ? 0 : 1));
            prev.swap(cur);
        }
        return (Real)prev[(size_t)n];
    }
};

using Edge = std::tuple<Index, Index, Real>;

template <class Atom, class Distance>
int systolic(const char* infile, Real radius, const char* outfile) {
    const SingleComm comm = COMM_SINGLE;
    PointContainer<Atom> mypoints;
    Distance distance;
    Index num_points;
    if constexpr (std::is_same<Atom, float>::value) num_points = mypoints.read_fvecs(infile, comm);
    else num_points = mypoints.read_seqs(infile, comm);
    CoverTree search(1.5, 10);
    search.build(mypoints, distance);
    std::vector<Edge> myedges;
    auto functor = [&](const Point<Atom>& p, const Point<Atom>& q, Real dist) { myedges.emplace_back(q.id(), p.id(), dist); };
    PointContainer<Atom> sendbuf = mypoints, recvbuf;
    typename PointContainer<Atom>::SendrecvRequest request;
    Index found = 0, neighbored = 0;
    const int nprocs = 1;
    for (int step = 0; step <= nprocs / 2; ++step) {
        sendbuf.sendrecv(recvbuf, 0, 0, comm, request);
        for (Index i = 0; i < sendbuf.num_points(); ++i) {
            found += search.radius_query(mypoints, distance, sendbuf[i], radius, functor);
            neighbored += search.has_radius_neighbor(mypoints, distance, sendbuf[i], radius) ? 1 : 0;
        }
        request.wait();
        sendbuf.swap(recvbuf);
    }
    Graph graph(myedges, num_points);
    graph.redistribute_edges(comm);
    graph.write_file(outfile, comm);
    std::printf("points %lld found %lld neighbored %lld edges %lld\n", (long long)num_points, (long long)found,
                (long long)neighbored, (long long)graph.num_edges());
    return 0;
}

int landmark(const char* infile, Real radius, Index num_centers, int seed, const char* outfile, bool served) {
    const SingleComm comm = COMM_SINGLE;
    PointContainer<float> mypoints;
    L2Distance<float> distance;
    const Index num_points = mypoints.read_fvecs(infile, comm);
    IndexVector landmarks;
    selection_sample(num_points, num_centers, landmarks, seed);
    PointContainer<float> mycenters, centers;
    mycenters.indexed_gather(mypoints, landmarks);
    centers.allgather(mycenters, comm);
    VoronoiDiagram<float> diagram(mypoints, centers, distance);
    std::vector<VoronoiCell<float>> mycells;
    diagram.coalesce_cells(mycells, comm);
    diagram.add_ghost_points_systolic(mycells, distance, radius, 1.5, 10, comm);
    std::vector<Edge> myedges;
    auto functor = [&](const Point<float>& p, const Point<float>& q, Real dist) { myedges.emplace_back(q.id(), p.id(), dist); };
    Index found = 0, ghosts = 0;
    for (const VoronoiCell<float>& cell : mycells) {
        CoverTree tree(1.5, 10);
        tree.build(cell, distance);
        ghosts += cell.num_ghosts();
        if (served) {
            found += tree.radius_query_batched(cell, distance, cell, radius, functor);
            found += tree.radius_query_batched(cell, distance, cell.ghosts(), radius, functor);
        } else {
            // the same queries through the plain container type: GPU joins
            const PointContainer<float>& members = cell;
            found += tree.radius_query_batched(members, distance, members, radius, functor);
            found += tree.radius_query_batched(members, distance, cell.ghosts(), radius, functor);
        }
    }
    Graph graph(myedges, num_points);
    graph.redistribute_edges(comm);
    graph.write_file(outfile, comm);
    std::printf("points %lld cells %zu ghosts %lld found %lld edges %lld\n", (long long)num_points, mycells.size(),
                (long long)ghosts, (long long)found, (long long)graph.num_edges());
    return 0;
}

int rips(const char* infile, Real radius, Index maxdim, Index num_centers, int seed, const char* prefix, bool use_ids) {
    const SingleComm comm = COMM_SINGLE;
    PointContainer<float> mypoints;
    L2Distance<float> distance;
    const Index num_points = mypoints.read_fvecs(infile, comm);
    IndexVector landmarks;
    selection_sample(num_points, num_centers, landmarks, seed);
    PointContainer<float> mycenters, centers;
    mycenters.indexed_gather(mypoints, landmarks);
    centers.allgather(mycenters, comm);
    VoronoiDiagram<float> diagram(mypoints, centers, distance);
    std::vector<VoronoiCell<float>> mycells;
    diagram.coalesce_cells(mycells, comm);
    diagram.add_ghost_points_systolic_rips(mycells, distance, radius, 1.5, 10, comm);
    Index interior = 0;
    for (const auto& cell : mycells)
        for (bool b : cell.interiors()) interior += b ? 1 : 0;
    std::vector<VoronoiComplex<float>> complexes;
    for (const VoronoiCell<float>& cell : mycells) {
        complexes.emplace_back(cell, radius, maxdim);
        complexes.back().build_filtration(distance, num_points, 1.5, 10);
    }
    std::vector<WeightedSimplex> mysimplices;
    Index num_simplices = 0;
    for (const auto& c : complexes) num_simplices += c.num_simplices();
    mysimplices.reserve((size_t)num_simplices);
    for (const auto& c : complexes) mysimplices.insert(mysimplices.end(), c.begin(), c.end());
    const std::string s = std::string(prefix) + ".txt";
    merge_and_write_filtration(s.c_str(), mysimplices, num_points, use_ids, comm);
    std::printf("points %lld cells %zu interior %lld simplices %lld\n", (long long)num_points, mycells.size(),
                (long long)interior, (long long)num_simplices);
    return 0;
}

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "systolic" && argc == 6) {
        const Real r = (Real)std::atof(argv[3]);
        if (std::string(argv[5]) == "edit") return systolic<char, EditDistance<char>>(argv[2], r, argv[4]);
        return systolic<float, L2Distance<float>>(argv[2], r, argv[4]);
    }
    if (mode == "landmark" && argc == 8)
        return landmark(argv[2], (Real)std::atof(argv[3]), std::atoll(argv[4]), std::atoi(argv[5]), argv[6],
                        std::string(argv[7]) == "served");
    if (mode == "rips" && argc == 9)
        return rips(argv[2], (Real)std::atof(argv[3]), std::atoll(argv[4]), std::atoll(argv[5]), std::atoi(argv[6]),
                    argv[7], std::string(argv[8]) == "ids");
    std::fprintf(stderr, "usage: see the header of tests/host/driver_bodies.cpp\n");
    return 2;
}
