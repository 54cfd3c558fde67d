// Vietoris-Rips filtration on one GPU: the rips driver's flags and merged
// output file "<outfile>.txt" (drivers/rips.cpp, include/simplex.hpp:130-196).
// The union of the per-cell complexes is the Rips complex of the whole point
// set, built here directly; maxdim up to 3 (tetrahedra for up to 55,111
// points, where the reference's simplex ids are exact).
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
    std::fprintf(stderr, "Options: -m INT   number of centers [%lld] (accepted; one GPU holds every cell)\n", (long long)num_centers);
    std::fprintf(stderr, "         -d INT   maximum simplex dimension [%lld] (0..3)\n", (long long)maxdim);
    std::fprintf(stderr, "         -c FLOAT cover tree base [%.2f] (accepted; no tree is built)\n", cover);
    std::fprintf(stderr, "         -l INT   leaf size [%lld] (accepted; no tree is built)\n", (long long)leaf_size);
    std::fprintf(stderr, "         -v INT   verbosity level [%d]\n", verbosity);
    std::fprintf(stderr, "         -D STR   distance metric [%s] (l2)\n", metric);
    std::fprintf(stderr, "         -o FILE  output filtration prefix (writes FILE.txt)\n");
    std::fprintf(stderr, "         -s INT   random number seed (accepted)\n");
    std::fprintf(stderr, "         -F       fix number of centers (accepted)\n");
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
    PointContainer<float> points;
    const Index n = points.read_file(infile);
    if (verbosity >= 1) std::fprintf(stderr, "[time=%.3f] read input file '%s' [size=%lld]\n", wall() - t0, infile, (long long)n);
    t = wall();
    rnng_filtration_t f = nullptr;
    rnng_host::check(rnng_rips_build(points.device_points(), radius, (int32_t)maxdim, &f, nullptr), "rips_build");
    int64_t ns = 0, per[4] = {0, 0, 0, 0};
    rnng_host::check(rnng_filtration_size(f, &ns, per), "filtration_size");
    std::vector<float> val((size_t)ns);
    std::vector<int64_t> ids((size_t)ns);
    rnng_host::check(rnng_filtration_copy(f, val.data(), ids.data(), ns), "filtration_copy");
    rnng_filtration_destroy(f);
    if (verbosity >= 1)
        std::fprintf(stderr, "[time=%.3f] built filtration [vertices=%lld,edges=%lld,triangles=%lld,tetrahedra=%lld]\n",
                     wall() - t, (long long)per[0], (long long)per[1], (long long)per[2], (long long)per[3]);
    std::fprintf(stderr, "[time=%.3f] complete\n", wall() - t0);
    if (outfile) {
        t = wall();
        std::string s = std::string(outfile) + ".txt";
        // already sorted by (value, id) and unique by id on the device
        FILE* fo = std::fopen(s.c_str(), "w");
        if (!fo) {
            std::fprintf(stderr, "cannot open %s\n", s.c_str());
            return 1;
        }
        for (int64_t k = 0; k < ns; ++k) {
            if (use_ids) std::fprintf(fo, "%f\t%lld\n", val[(size_t)k], (long long)ids[(size_t)k]);
            else std::fprintf(fo, "%f\t%s\n", val[(size_t)k], Simplex(ids[(size_t)k]).get_simplex_repr(n).c_str());
        }
        std::fclose(fo);
        if (verbosity >= 1) std::printf("[time=%.3f] wrote merged simplices to '%s'\n", wall() - t, s.c_str());
    }
    return 0;
}
