// C++ host layer over the rnng_amd C ABI: the reference's header API
// (include/utils.h, point.h, search.h, graph.h, simplex.h) for one process
// and one GPU.  Names, argument meanings and error behaviour follow the
// reference: a failed check prints the library's message and aborts, as the
// reference's always-on asserts do.  Calls that take an MPI communicator in
// the reference are provided in their single-rank, communicator-free form.
#ifndef RNNG_HOST_UTILS_H_
#define RNNG_HOST_UTILS_H_

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../rnng_amd.h"

using Index = int64_t;                 // include/utils.h:24
using Real = float;                    // include/utils.h:25
using IndexVector = std::vector<Index>;
using RealVector = std::vector<Real>;
using BoolVector = std::vector<bool>;

// The distance functors are the drivers' own (drivers/systolic.cpp:31-41);
// the host layer recognises them by type and runs their GPU kernels.
template <class Atom> struct L2Distance;
template <class Atom> struct EditDistance;

namespace rnng_host {

[[noreturn]] inline void fail(const std::string& msg) {
    std::fprintf(stderr, "rnng: %s\n", msg.c_str());
    std::fflush(stderr);
    std::abort();
}

inline void check(int rc, const char* what) {
    if (rc != RNNG_OK) {
        std::fprintf(stderr, "rnng: %s failed (%d): %s\n", what, rc, rnng_last_error());
        std::fflush(stderr);
        std::abort();
    }
}

template <class Distance>
constexpr bool is_l2() {
    return std::is_same<std::remove_cv_t<Distance>, L2Distance<float>>::value;
}
template <class Distance>
constexpr bool is_edit() {
    return std::is_same<std::remove_cv_t<Distance>, EditDistance<char>>::value;
}

}  // namespace rnng_host

// One process's communicator.  The reference passes an MPI_Comm to every call
// that moves data between ranks (reads, the systolic ring, cell exchange,
// edge redistribution, the filtration merge); with one process those calls
// take this token, keep the reference's argument lists and move nothing.
struct SingleComm {};
inline constexpr SingleComm COMM_SINGLE{};

// Knuth's selection sampling with the reference's engine and distribution
// (include/utils.hpp:1-29): `size` ascending ids from [0, range).
inline void selection_sample(Index range, Index size, IndexVector& sample, int seed) {
    sample.resize((size_t)size);
    rnng_host::check(rnng_selection_sample(range, size, sample.data(), seed), "selection_sample");
}

#endif
