// Right-hand-side generator of the reference's benchmark driver, for the
// tests that compare iteration counts with the reference's published result
// cells (test infrastructure only). benchmarks/src/benchmark.cpp:440-455
// (make_rhs) draws g[i] ~ N(0, 1) for i = 0 .. n-1, in order, from one
// std::normal_distribution<double>(0, 1) fed by std::mt19937_64(seed); the
// caller then forms b = L g, centres it on a Laplacian and normalises it. This
// file draws g exactly so (the standard library's engine and distribution,
// as the reference's build uses them); the rest is done by the test in numpy.
//
// Built by tests/golden/Makefile into tests/golden/librefrhs.so.
#include <cstdint>
#include <random>

extern "C" {

// g[0..n) as make_rhs draws it.
void refrhs_normal(std::uint64_t seed, std::int64_t n, double* g) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> N(0.0, 1.0);
    for (std::int64_t i = 0; i < n; ++i) g[i] = N(rng);
}

// The 10000th output of a default-constructed std::mt19937_64, which the C++
// standard fixes at 9981545732273789042 ([rand.predef]): pins the engine.
std::uint64_t refrhs_mt64_10000th() {
    std::mt19937_64 rng;
    rng.discard(9999);
    return rng();
}

}  // extern "C"
