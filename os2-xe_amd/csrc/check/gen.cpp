// gen.cpp - input matrices of the accuracy harness (oz2/bench_accuracy.py), generated
// the way the reference's accuracy benchmark generates them
// (benchmarks/bench_accuracy.cpp:41-69), so the same options give the same inputs:
// std::mt19937_64 seeded per matrix, the standard uniform [0, 1) and normal (0, 1)
// distributions, column-major rows x cols.  Measurement harness only; not part of the
// product library.
//
// Distributions: "urand" 2U - 1; "phiX" (U - 0.5) exp(X N) (one uniform draw, then one
// normal draw, per entry); "sincos" sin(r + c) for A, cos(r c) for B; "spike" urand
// with one entry per row of A scaled by 2^40; "cancel" urand with the second half of
// A's columns repeating the first and the second half of B's rows cancelling the first
// up to 2^-30 of its own value.
#include <cmath>
#include <cstdint>
#include <random>
#include <string>

extern "C" int oz2chk_gen(const char* dist, int64_t rows, int64_t cols, uint64_t seed, int is_a, double* X) {
    if (!dist || !X || rows < 0 || cols < 0) return 1;
    const std::string d(dist);
    const bool uniform = d == "urand" || d == "spike" || d == "cancel";
    const bool lognormal = d.rfind("phi", 0) == 0 && d.size() > 3;
    if (!uniform && !lognormal && d != "sincos") return 1;
    double phi = 0.0;
    if (lognormal) {
        try {
            phi = std::stod(d.substr(3));
        } catch (...) {
            return 1;
        }
    }
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    std::normal_distribution<double> gauss(0.0, 1.0);
    for (int64_t c = 0; c < cols; ++c)
        for (int64_t r = 0; r < rows; ++r) {
            double v;
            if (uniform) {
                v = 2.0 * unif(rng) - 1.0;
            } else if (!lognormal) {
                v = is_a ? std::sin(static_cast<double>(r + c)) : std::cos(static_cast<double>(r) * static_cast<double>(c));
            } else {
                const double u = unif(rng);  // the uniform draw comes first
                const double g = gauss(rng);
                v = (u - 0.5) * std::exp(phi * g);
            }
            X[r + c * rows] = v;
        }
    if (d == "spike" && is_a && cols > 0)
        for (int64_t r = 0; r < rows; ++r) X[r + ((r * 7919) % cols) * rows] *= std::ldexp(1.0, 40);
    if (d == "cancel") {
        if (is_a) {
            const int64_t h = cols / 2;
            for (int64_t c = h; c < 2 * h; ++c)
                for (int64_t r = 0; r < rows; ++r) X[r + c * rows] = X[r + (c - h) * rows];
        } else {
            const int64_t h = rows / 2;
            for (int64_t c = 0; c < cols; ++c)
                for (int64_t r = h; r < 2 * h; ++r)
                    X[r + c * rows] = -X[(r - h) + c * rows] + std::ldexp(X[r + c * rows], -30);
        }
    }
    return 0;
}
