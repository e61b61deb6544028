// gen.cpp -- synthetic inputs for the hot-path workloads (host code).
//
// * vvhd_gen_wake: the Karman-street wake of the 1M / 4M single-step
//   benchmarks: two staggered rows of 200 Gaussian clusters each.
// * vvhd_gen_cylinder: corners of a circle discretised into N equal
//   segments, as gen_arc_N (utils/vvflow/gen_body.cpp:74-88) called by
//   gen_cylinder (utils/vvflow/gen_cylinder.cpp:34-38) produces them.

#include "vvhd_gpu.h"

#include <cmath>
#include <cstdint>

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u53() { return (double)(next() >> 11) * 0x1.0p-53; }   // [0, 1)
};

}  // namespace

extern "C" int vvhd_gen_wake(int64_t n, uint64_t seed, double* objs6) {
    if (n < 0 || (n > 0 && !objs6)) return VVHD_EINVAL;
    const int kClusters = 400;
    const double kPi = 3.14159265358979323846;
    SplitMix64 rng{seed};
    int64_t w = 0;
    for (int c = 0; c < kClusters; c++) {
        const bool top = c < kClusters / 2;
        const int k = top ? c : c - kClusters / 2;
        const double cx = top ? 0.3 * k : 0.15 + 0.3 * k;
        const double cy = top ? 0.2 : -0.2;
        const double sign = top ? -1.0 : 1.0;
        const int64_t count = n / kClusters + (c < n % kClusters ? 1 : 0);
        for (int64_t b = 0; b < count; b++) {
            const double u1 = 1.0 - rng.u53();   // (0, 1]
            const double u2 = rng.u53();
            const double r = std::sqrt(-2.0 * std::log(u1));
            const double z1 = r * std::cos(2.0 * kPi * u2);
            const double z2 = r * std::sin(2.0 * kPi * u2);
            double* o = objs6 + 6 * w++;
            o[0] = cx + 0.05 * z1;
            o[1] = cy + 0.05 * z2;
            o[2] = sign * (0.5 + rng.u53()) * 1e-3;
            o[3] = 0.0;
            o[4] = 0.0;
            o[5] = 0.0;
        }
    }
    return VVHD_OK;
}

extern "C" int vvhd_gen_cylinder(double radius, int32_t n, double* corners_xy) {
    if (n <= 0 || !(radius > 0) || !corners_xy) return VVHD_EINVAL;
    const double a1 = 2 * M_PI, a2 = 0;
    for (int32_t i = 0; i < n; i++) {
        const double ai = a1 + (a2 - a1) * double(i) / double(n);
        corners_xy[2 * i] = 0 + radius * std::cos(ai);
        corners_xy[2 * i + 1] = 0 + radius * std::sin(ai);
    }
    return VVHD_OK;
}
