// Host-only checks of the C++ layer (no GPU call): simplex ids and vertex
// lists, the graph writers and the fvecs / seqs readers.  Driven by
// tests/test_host_cpu.py, which compares the printed results with values it
// computes independently.
#include <iostream>

#include "rnng/graph.h"
#include "rnng/simplex.h"

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "simplex") {
        // lines of vertices on stdin -> "id repr"
        Index k;
        while (std::cin >> k) {
            IndexVector v((size_t)k);
            for (auto& x : v) std::cin >> x;
            Simplex s(v);
            std::cout << s.getid() << " " << s.getdim() << " " << s.get_simplex_repr(1000000) << "\n";
        }
        return 0;
    }
    if (mode == "graph") {
        // "n m" then m lines "u v w" -> redistribute, write both formats
        Index n, m;
        std::cin >> n >> m;
        Graph::EdgeVector e;
        for (Index i = 0; i < m; ++i) {
            Index u, v;
            Real w;
            std::cin >> u >> v >> w;
            e.emplace_back(u, v, w);
        }
        Graph g(e, n);
        g.redistribute_edges();
        g.write_file(argv[2]);
        g.write_file_stream(argv[3]);
        return 0;
    }
    if (mode == "fvecs" || mode == "file") {
        PointContainer<float> P;
        const Index n = mode == "file" ? P.read_file(argv[2]) : P.read_fvecs(argv[2]);
        std::cout << n << " " << P.dim() << "\n";
        for (Index i = 0; i < n; ++i) {
            const auto p = P[i];
            std::cout << p.id();
            for (Index k = 0; k < p.size(); ++k) std::cout << " " << p[k];
            std::cout << "\n";
        }
        return 0;
    }
    if (mode == "seqs") {
        PointContainer<char> S;
        const Index n = S.read_seqs(argv[2]);
        std::cout << n << "\n";
        for (Index i = 0; i < n; ++i) {
            const auto p = S[i];
            std::cout << p.id() << " " << std::string(p.begin(), p.end()) << "\n";
        }
        return 0;
    }
    return 2;
}
