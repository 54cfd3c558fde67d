// Shared pieces of the command-line drivers: the distance functors the
// reference's drivers define (drivers/systolic.cpp:253-299) and timing lines
// in the drivers' "[time=...]" form.
#pragma once

#include <getopt.h>
#include <sys/time.h>

#include <cmath>

#include <vector>

#include "../../include/rnng/graph.h"
#include "../../include/rnng/search.h"

// Sequential FP32 sum of (p_i - q_i)^2, then sqrt: the functor the GPU
// kernels reproduce bit for bit.
template <class Atom>
struct L2Distance {
    Real operator()(const Point<Atom>& p, const Point<Atom>& q) const {
        volatile Real acc = 0;   // no contraction, index order
        for (Index i = 0; i < p.size(); ++i) {
            const Real t = p[i] - q[i];
            const Real t2 = t * t;
            acc = acc + t2;
        }
        return std::sqrt((Real)acc);
    }
};

// Levenshtein distance with unit costs.
template <class Atom>
struct EditDistance {
    Real operator()(const Point<Atom>& s, const Point<Atom>& t) const {
        const Index m = s.size(), n = t.size();
        std::vector<Index> prev((size_t)n + 1), cur((size_t)n + 1);
        for (Index j = 0; j <= n; ++j) prev[(size_t)j] = j;
        for (Index i = 1; i <= m; ++i) {
            cur[0] = i;
            for (Index j = 1; j <= n; ++j) {
                const Index sub = prev[(size_t)j - 1] + (s[i - 1] == t[j - 1] ? 0 : 1);
                cur[(size_t)j] = std::min(std::min(prev[(size_t)j] + 1, cur[(size_t)j - 1] + 1), sub);
            }
            prev.swap(cur);
        }
        return (Real)prev[(size_t)n];
    }
};

inline double wall() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}
