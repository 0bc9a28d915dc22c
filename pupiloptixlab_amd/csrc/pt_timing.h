// Kernel-time bookkeeping shared by the library's event collection and the
// host tests (tests/timing_sim.cpp).
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

namespace pupil {

// Total length of the union of [start, end) intervals (the time during which
// at least one of them is open); sorts `iv`.
inline double interval_union(std::vector<std::pair<float, float>> &iv) {
    if (iv.empty()) return 0.0;
    std::sort(iv.begin(), iv.end());
    double total = 0.0;
    float lo = iv.front().first, hi = iv.front().second;
    for (const auto &x : iv) {
        if (x.first > hi) {
            total += hi - lo;
            lo = x.first;
            hi = x.second;
        } else {
            hi = std::max(hi, x.second);
        }
    }
    return total + (hi - lo);
}

}  // namespace pupil
