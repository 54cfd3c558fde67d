// Epsilon graph of a whole point set on one GPU: the systolic driver's flags
// and output (drivers/systolic.cpp), the ring over ranks replaced by one
// self join.  -D l2 reads fvecs, -D edit reads one string per line.
#include "cli_common.h"

static Real radius = -1;
static const char* infile = nullptr;
static const char* outfile = nullptr;
static const char* metric = "l2";
static Real cover = 1.5;
static Index leaf_size = 10;
static int verbosity = 1;

static void usage(const char* prog, int err) {
    std::fprintf(stderr, "Usage: %s [options] -i <points> -r <radius>\n", prog);
    std::fprintf(stderr, "Options: -c FLOAT cover tree base [%.2f] (accepted; no tree is built)\n", cover);
    std::fprintf(stderr, "         -l INT   leaf size [%lld] (accepted; no tree is built)\n", (long long)leaf_size);
    std::fprintf(stderr, "         -v INT   verbosity level [%d]\n", verbosity);
    std::fprintf(stderr, "         -D STR   metric [%s]\n", metric);
    std::fprintf(stderr, "         -o FILE  output edge file\n");
    std::fprintf(stderr, "         -h       help message\n");
    std::exit(err);
}

template <class Atom>
static int run() {
    double t0 = wall(), t;
    PointContainer<Atom> points;
    Index n;
    if constexpr (std::is_same<Atom, float>::value) n = points.read_file(infile);
    else n = points.read_seqs(infile);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] read input file '%s' [size=%lld]\n", wall() - t0, infile, (long long)n);
    t = wall();
    rnng_host::GraphOwner g;
    rnng_stats st{};
    if constexpr (std::is_same<Atom, float>::value) {
        const int32_t mode = n > 0 && points[0].size() <= 3 ? RNNG_MODE_GRID : RNNG_MODE_MFMA;
        rnng_host::check(rnng_graph_build(points.device_points(), radius, mode, &g.h, &st), "graph_build");
    }
    else
        rnng_host::check(rnng_edit_graph_build(points.device_strings(), radius, &g.h, &st), "edit_graph_build");
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] found neighbors\n", wall() - t);
    t = wall();
    Graph graph = Graph::from_device(g.h, n);
    if (verbosity >= 1)
        std::fprintf(stderr, "[time=%.3f] redistributed edges [points=%lld,edges=%lld,density=%.3f]\n", wall() - t,
                     (long long)n, (long long)graph.num_edges(), (graph.num_edges() + 0.0) / (double)std::max<Index>(n, 1));
    if (outfile) {
        t = wall();
        graph.write_file_stream(outfile);
        if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] wrote edges to file '%s'\n", wall() - t, outfile);
    }
    std::fprintf(stderr, "[time=%.3f] complete\n", wall() - t0);
    return 0;
}

int main(int argc, char* argv[]) {
    int c;
    while ((c = getopt(argc, argv, "i:r:c:l:v:o:D:h")) >= 0) {
        if (c == 'i') infile = optarg;
        else if (c == 'r') radius = (Real)atof(optarg);
        else if (c == 'c') cover = (Real)atof(optarg);
        else if (c == 'l') leaf_size = atoi(optarg);
        else if (c == 'v') verbosity = atoi(optarg);
        else if (c == 'D') metric = optarg;
        else if (c == 'o') outfile = optarg;
        else if (c == 'h') usage(argv[0], 0);
        else usage(argv[0], 1);
    }
    if (!infile || radius < 0) usage(argv[0], 1);
    if (!std::strcmp(metric, "l2")) return run<float>();
    if (!std::strcmp(metric, "edit")) return run<char>();
    std::fprintf(stderr, "unknown metric '%s'\n", metric);
    return 1;
}
