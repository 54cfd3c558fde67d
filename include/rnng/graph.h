// The epsilon graph as the drivers write it (include/graph.h, graph.hpp):
// weighted edges sorted by (u, v, w); on one rank redistribute_edges makes the
// list symmetric and unique.
#ifndef RNNG_HOST_GRAPH_H_
#define RNNG_HOST_GRAPH_H_

#include <sstream>
#include <iomanip>

#include "point.h"

class Graph {
public:
    using Edge = std::tuple<Index, Index, Real>;
    using EdgeVector = std::vector<Edge>;

    Graph(const EdgeVector& edges, Index num_vertices) : edges(edges), num_vertices(num_vertices) {
        std::sort(this->edges.begin(), this->edges.end());
    }

    // The directed edge list the drivers hold after redistribute_edges for a
    // self-join graph on the GPU: both orientations of every edge and a
    // (v, v, 0) loop per vertex, sorted on the device.
    static Graph from_device(rnng_graph_t g, Index num_vertices) {
        int64_t need = 0;
        rnng_host::check(rnng_graph_copy_directed(g, num_vertices, nullptr, nullptr, 0, &need), "graph_copy_directed");
        std::vector<int64_t> uv((size_t)(2 * need));
        std::vector<float> w((size_t)need);
        rnng_host::check(rnng_graph_copy_directed(g, num_vertices, uv.data(), w.data(), need, &need),
                         "graph_copy_directed");
        Graph out;
        out.num_vertices = num_vertices;
        out.edges.resize((size_t)need);
        for (int64_t e = 0; e < need; ++e) out.edges[(size_t)e] = Edge(uv[2 * e], uv[2 * e + 1], w[(size_t)e]);
        return out;   // already in (u, v) order, one weight per (u, v)
    }

    // One rank of include/graph.hpp:59-116: every edge to the owner of both
    // endpoints (both orientations), sorted and unique.
    void redistribute_edges() {
        EdgeVector all;
        all.reserve(2 * edges.size());
        for (const auto& [u, v, w] : edges) {
            all.emplace_back(u, v, w);
            all.emplace_back(v, u, w);
        }
        std::sort(all.begin(), all.end());
        all.erase(std::unique(all.begin(), all.end()), all.end());
        edges.swap(all);
    }

    void redistribute_edges(SingleComm) { redistribute_edges(); }

    // "% n n m" then "u v w" per edge, w as %.6f (include/graph.hpp:1-18).
    void write_file(const char* fname) const {
        FILE* f = std::fopen(fname, "w");
        if (!f) fail(fname);
        std::fprintf(f, "%% %lld %lld %lld\n", (long long)num_vertices, (long long)num_vertices,
                     (long long)edges.size());
        for (const auto& [i, j, w] : edges) std::fprintf(f, "%lld %lld %.6f\n", (long long)i, (long long)j, w);
        std::fclose(f);
    }

    // The file the drivers write (the communicator form, include/graph.hpp:20-57,
    // on one rank): w through an ostream at setprecision(6).
    void write_file_stream(const char* fname) const {
        FILE* f = std::fopen(fname, "w");
        if (!f) fail(fname);
        std::ostringstream ss;
        ss << "% " << num_vertices << " " << num_vertices << " " << (Index)edges.size() << "\n";
        ss << std::setprecision(6);
        const size_t flush_at = 1 << 24;
        for (const auto& [u, v, w] : edges) {
            ss << u << " " << v << " " << w << "\n";
            if (ss.tellp() > (std::streamoff)flush_at) {
                const std::string s = ss.str();
                std::fwrite(s.data(), 1, s.size(), f);
                ss.str("");
                ss.clear();
            }
        }
        const std::string s = ss.str();
        std::fwrite(s.data(), 1, s.size(), f);
        std::fclose(f);
    }

    // The communicator form the drivers call (include/graph.hpp:20-57).
    void write_file(const char* fname, SingleComm) const { write_file_stream(fname); }

    Index num_edges() const { return (Index)edges.size(); }
    const EdgeVector& get_edges() const { return edges; }

private:
    Graph() : num_vertices(0) {}
    [[noreturn]] static void fail(const char* fname) {
        std::fprintf(stderr, "rnng: cannot open %s\n", fname);
        std::abort();
    }
    EdgeVector edges;
    Index num_vertices;
};

#endif
