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

// The systolic driver's sequence (drivers/systolic.cpp:70-195) on the class
// API: read, build the search structure, query the rank's block, build the
// graph and redistribute it, write.  With one rank the ring has one step, in
// which the rank queries its own block, so the queries run as one batched self
// query.
template <class Atom, class Distance>
static int run() {
    double t0 = wall(), t;
    PointContainer<Atom> mypoints;
    Distance distance;
    Index num_points;
    if constexpr (std::is_same<Atom, float>::value) num_points = mypoints.read_file(infile);
    else num_points = mypoints.read_seqs(infile, COMM_SINGLE);
    if (verbosity >= 1)
        std::fprintf(stderr, "[time=%.3f] read input file '%s' [size=%lld]\n", wall() - t0, infile, (long long)num_points);
    t = wall();
    CoverTree search(cover, leaf_size);
    search.build(mypoints, distance);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] built cover tree\n", wall() - t);
    t = wall();
    using Edge = std::tuple<Index, Index, Real>;
    std::vector<Edge> myedges;
    auto functor = [&](const Point<Atom>& p, const Point<Atom>& q, Real dist) { myedges.emplace_back(q.id(), p.id(), dist); };
    search.radius_query_batched(mypoints, distance, mypoints, radius, functor);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] found neighbors\n", wall() - t);
    t = wall();
    Graph graph(myedges, num_points);
    graph.redistribute_edges(COMM_SINGLE);
    if (verbosity >= 1)
        std::fprintf(stderr, "[time=%.3f] redistributed edges [points=%lld,edges=%lld,density=%.3f]\n", wall() - t,
                     (long long)num_points, (long long)graph.num_edges(),
                     (graph.num_edges() + 0.0) / (double)std::max<Index>(num_points, 1));
    if (outfile) {
        t = wall();
        graph.write_file(outfile, COMM_SINGLE);
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
    if (!std::strcmp(metric, "l2")) return run<float, L2Distance<float>>();
    if (!std::strcmp(metric, "edit")) return run<char, EditDistance<char>>();
    std::fprintf(stderr, "unknown metric '%s'\n", metric);
    return 1;
}
