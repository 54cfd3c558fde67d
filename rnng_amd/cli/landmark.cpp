// Epsilon graph through Voronoi landmark cells on one GPU: the landmark
// driver's flags and output (drivers/landmark.cpp): selection-sampled centers,
// Voronoi diagram, cells, ghosts, per-cell queries, the symmetric edge file.
#include "cli_common.h"

static Real radius = -1;
static const char* infile = nullptr;
static const char* outfile = nullptr;
static const char* metric = "l2";
static Real cover = 1.5;
static Index leaf_size = 10;
static Index num_centers = 1;
static int rng_seed = -1;
static int verbosity = 1;

static void usage(const char* prog, int err) {
    std::fprintf(stderr, "Usage: %s [options] -i <points> -r <radius>\n", prog);
    std::fprintf(stderr, "Options: -m INT   number of centers [%lld]\n", (long long)num_centers);
    std::fprintf(stderr, "         -c FLOAT cover tree base [%.2f] (accepted; no tree is built)\n", cover);
    std::fprintf(stderr, "         -l INT   leaf size [%lld] (accepted; no tree is built)\n", (long long)leaf_size);
    std::fprintf(stderr, "         -v INT   verbosity level [%d]\n", verbosity);
    std::fprintf(stderr, "         -D STR   distance metric [%s] (l2)\n", metric);
    std::fprintf(stderr, "         -o FILE  output edge file\n");
    std::fprintf(stderr, "         -s INT   random number seed\n");
    std::fprintf(stderr, "         -F       fix number of centers (one rank: always fixed)\n");
    std::fprintf(stderr, "         -h       help message\n");
    std::exit(err);
}

int main(int argc, char* argv[]) {
    int c;
    while ((c = getopt(argc, argv, "i:r:m:c:l:v:o:s:FD:h")) >= 0) {
        if (c == 'i') infile = optarg;
        else if (c == 'r') radius = (Real)atof(optarg);
        else if (c == 'c') cover = (Real)atof(optarg);
        else if (c == 'l') leaf_size = atoi(optarg);
        else if (c == 'v') verbosity = atoi(optarg);
        else if (c == 'm') num_centers = atoi(optarg);
        else if (c == 'o') outfile = optarg;
        else if (c == 'F') {
        } else if (c == 's') rng_seed = atoi(optarg);
        else if (c == 'D') metric = optarg;
        else if (c == 'h') usage(argv[0], 0);
        else usage(argv[0], 1);
    }
    if (!infile || radius < 0) usage(argv[0], 1);
    if (std::strcmp(metric, "l2") != 0) {
        std::fprintf(stderr, "the GPU landmark path computes the l2 metric\n");
        return 1;
    }
    double t0 = wall(), t;
    // the landmark driver's sequence (drivers/landmark.cpp:70-250) on the class API
    PointContainer<float> mypoints;
    L2Distance<float> distance;
    const Index num_points = mypoints.read_file(infile);
    if (verbosity >= 1)
        std::fprintf(stderr, "[time=%.3f] read input file '%s' [size=%lld]\n", wall() - t0, infile, (long long)num_points);
    if (num_centers > num_points || num_centers < 1) {
        std::fprintf(stderr, "error: need 1 <= centers <= points (-m)\n");
        return 1;
    }
    t = wall();
    IndexVector landmarks;
    selection_sample(num_points, num_centers, landmarks, rng_seed);
    PointContainer<float> mycenters, centers;
    mycenters.indexed_gather(mypoints, landmarks);
    centers.allgather(mycenters, COMM_SINGLE);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] picked centers [centers=%lld]\n", wall() - t, (long long)num_centers);
    t = wall();
    VoronoiDiagram<float> diagram(mypoints, centers, distance);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] computed point partitioning\n", wall() - t);
    t = wall();
    std::vector<VoronoiCell<float>> mycells;
    diagram.coalesce_cells(mycells, COMM_SINGLE);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] coalesced voronoi cells\n", wall() - t);
    t = wall();
    diagram.add_ghost_points_systolic(mycells, distance, radius, cover, leaf_size, COMM_SINGLE);
    Index ghosts = 0;
    for (const auto& cell : mycells) ghosts += cell.num_ghosts();
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] added ghost points [ghosts=%lld]\n", wall() - t, (long long)ghosts);
    t = wall();
    using Edge = std::tuple<Index, Index, Real>;
    std::vector<Edge> myedges;
    auto functor = [&](const Point<float>& p, const Point<float>& q, Real dist) { myedges.emplace_back(q.id(), p.id(), dist); };
    for (const VoronoiCell<float>& cell : mycells) {
        CoverTree tree(cover, leaf_size);
        tree.build(cell, distance);
        tree.radius_query_batched(cell, distance, cell, radius, functor);
        tree.radius_query_batched(cell, distance, cell.ghosts(), radius, functor);
    }
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
