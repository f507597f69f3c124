// The reference's summary of repeated timings (src/unixbench_score.cpp:9-20):
// the geometric mean of the fastest two thirds of the samples.  The samples
// are sorted in place and the slow third dropped, as in the reference; of a
// single sample nothing is kept and the result is 0/0 (NaN) there too.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace fdm {

inline double unixbench_score(std::vector<double>& data) {
    if (data.empty()) return 0.0;
    std::sort(data.begin(), data.end());
    const int keep = static_cast<int>((2 * data.size()) / 3);
    data.resize(keep);
    double log_sum = 0.0;
    for (double t : data) log_sum += std::log(t);
    return std::exp(log_sum / keep);
}

}  // namespace fdm
