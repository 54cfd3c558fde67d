// Vietoris-Rips filtration on one GPU: the rips driver's flags, sequence and
// output files (drivers/rips.cpp, include/point.hpp:1126-1290,
// include/simplex.hpp:130-196): landmark cells with ghosts, one complex per
// cell built on the GPU, the merged "<outfile>.txt"; maxdim up to 3
// (tetrahedra for up to 55,111 points, where the reference's simplex ids are
// exact).
#include "cli_common.h"

static Real radius = -1;
static const char* infile = nullptr;
static const char* outfile = nullptr;
static const char* metric = "l2";
static Real cover = 1.5;
static Index leaf_size = 10;
static Index num_centers = 1;
static Index maxdim = 3;
static int rng_seed = -1;
static int verbosity = 1;
static bool use_ids = true;

static void usage(const char* prog, int err) {
    std::fprintf(stderr, "Usage: %s [options] -i <points> -r <radius>\n", prog);
    std::fprintf(stderr, "Options: -m INT   number of centers [%lld]\n", (long long)num_centers);
    std::fprintf(stderr, "         -d INT   maximum simplex dimension [%lld] (0..3)\n", (long long)maxdim);
    std::fprintf(stderr, "         -c FLOAT cover tree base [%.2f] (accepted; no tree is built)\n", cover);
    std::fprintf(stderr, "         -l INT   leaf size [%lld] (accepted; no tree is built)\n", (long long)leaf_size);
    std::fprintf(stderr, "         -v INT   verbosity level [%d]\n", verbosity);
    std::fprintf(stderr, "         -D STR   distance metric [%s] (l2)\n", metric);
    std::fprintf(stderr, "         -o FILE  output filtration prefix (writes FILE.txt)\n");
    std::fprintf(stderr, "         -s INT   random number seed\n");
    std::fprintf(stderr, "         -F       fix number of centers (one rank: always fixed)\n");
    std::fprintf(stderr, "         -V       write vertex lists instead of simplex ids\n");
    std::fprintf(stderr, "         -h       help message\n");
    std::exit(err);
}

int main(int argc, char* argv[]) {
    int c;
    while ((c = getopt(argc, argv, "i:r:m:c:l:v:o:s:FD:d:Vh")) >= 0) {
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
        else if (c == 'd') maxdim = atoi(optarg);
        else if (c == 'V') use_ids = false;
        else if (c == 'h') usage(argv[0], 0);
        else usage(argv[0], 1);
    }
    if (!infile || radius < 0) usage(argv[0], 1);
    if (std::strcmp(metric, "l2") != 0 || maxdim < 0 || maxdim > 3) {
        std::fprintf(stderr, "the GPU Rips path computes the l2 metric up to dimension 3 (use -d 0..3)\n");
        return 1;
    }
    double t0 = wall(), t;
    // the rips driver's sequence (drivers/rips.cpp:70-250) on the class API
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
    diagram.add_ghost_points_systolic_rips(mycells, distance, radius, cover, leaf_size, COMM_SINGLE);
    if (verbosity >= 1) std::printf("[time=%.3f] added ghost points\n", wall() - t);
    t = wall();
    std::vector<VoronoiComplex<float>> complexes;
    for (const VoronoiCell<float>& cell : mycells) {
        complexes.emplace_back(cell, radius, maxdim);
        complexes.back().build_filtration(distance, num_points, cover, leaf_size);
    }
    if (verbosity >= 1) std::printf("[time=%.3f] built complexes\n", wall() - t);
    if (outfile) {
        // per-cell files: the driver passes use_ids where write_filtration_file
        // takes the vertex count (drivers/rips.cpp:204), so they hold ids
        t = wall();
        for (size_t i = 0; i < complexes.size(); ++i) {
            const std::string s = std::string(outfile) + ".rank0.cell" + std::to_string(i) + ".txt";
            complexes[i].write_filtration_file(s.c_str(), use_ids);
        }
        if (verbosity >= 1) std::printf("[time=%.3f] wrote complex files\n", wall() - t);
    }
    std::fprintf(stderr, "[time=%.3f] complete\n", wall() - t0);
    if (outfile) {
        t = wall();
        std::vector<WeightedSimplex> mysimplices;
        Index total = 0;
        for (const auto& c : complexes) total += c.num_simplices();
        mysimplices.reserve((size_t)total);
        for (const auto& c : complexes) mysimplices.insert(mysimplices.end(), c.begin(), c.end());
        const std::string s = std::string(outfile) + ".txt";
        merge_and_write_filtration(s.c_str(), mysimplices, num_points, use_ids, COMM_SINGLE);
        if (verbosity >= 1) std::printf("[time=%.3f] wrote merged simplices to '%s'\n", wall() - t, s.c_str());
    }
    return 0;
}
