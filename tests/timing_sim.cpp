// Host check of the kernel active-time rule (pupiloptixlab_amd/csrc/pt_timing.h):
// interval_union against a brute-force count over a fine time grid, on fixed
// cases (disjoint, nested, overlapping, touching, identical intervals) and on
// seeded random sets.  Prints "ok" or the first failure.
#include <cmath>
#include <cstdio>
#include <random>

#include "pt_timing.h"

using namespace pupil;

// Length covered by the intervals, counted on a grid of step 1/64 (the
// intervals below start and end on that grid).
static double brute(const std::vector<std::pair<float, float>> &iv) {
    double covered = 0.0;
    for (int t = 0; t < 64 * 200; ++t) {
        const double a = t / 64.0 + 1.0 / 128.0;
        for (const auto &x : iv)
            if (a > x.first && a < x.second) {
                covered += 1.0 / 64.0;
                break;
            }
    }
    return covered;
}

static bool same(std::vector<std::pair<float, float>> iv, double want) {
    const double got = interval_union(iv);
    if (std::fabs(got - want) > 1e-6) {
        std::printf("union %g, expected %g\n", got, want);
        return false;
    }
    return true;
}

int main() {
    if (!same({}, 0.0) || !same({{1.f, 3.f}}, 2.0) || !same({{1.f, 3.f}, {5.f, 6.f}}, 3.0) ||
        !same({{1.f, 10.f}, {2.f, 3.f}}, 9.0) || !same({{1.f, 4.f}, {3.f, 6.f}}, 5.0) ||
        !same({{1.f, 2.f}, {2.f, 3.f}}, 2.0) || !same({{2.f, 5.f}, {2.f, 5.f}, {2.f, 5.f}}, 3.0) ||
        !same({{5.f, 6.f}, {1.f, 2.f}, {1.5f, 5.5f}}, 5.0))
        return 1;
    std::mt19937 rng(7);
    for (int trial = 0; trial < 200; ++trial) {
        std::vector<std::pair<float, float>> iv;
        const int n = 1 + (int)(rng() % 12);
        for (int k = 0; k < n; ++k) {
            const int a = (int)(rng() % (64 * 180)), len = 1 + (int)(rng() % (64 * 20));
            iv.emplace_back(a / 64.f, (a + len) / 64.f);
        }
        const double want = brute(iv);
        if (!same(iv, want)) {
            std::printf("trial %d\n", trial);
            return 1;
        }
    }
    std::printf("ok\n");
    return 0;
}
