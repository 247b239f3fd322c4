// Host check of wrap_period (treecorr_amd/csrc/tc_common.h) against the reference's
// Position::wrap loop (include/Position.h:106-112): identical inside |d| <= 1.5 p, and
// bounded (a result in [-p/2, p/2] up to rounding) for tiny periods and far separations.
#include "tc_common.h"

#include <cmath>
#include <cstdio>
#include <random>

static double ref_wrap(double d, double p) {
    while (d > p / 2.) d -= p;
    while (d < -p / 2.) d += p;
    return d;
}

int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> u(-1.5, 1.5);
    long bad = 0, n = 0;
    const double periods[] = {1., 3.7, 1000., 1e-3, 360.};
    for (double p : periods) {
        for (int i = 0; i < 200000; ++i, ++n) {
            const double d = u(rng) * p;
            if (tc::wrap_period(d, p) != ref_wrap(d, p)) ++bad;
        }
        // exact boundary values
        const double edge[] = {p / 2., -p / 2., 1.5 * p, -1.5 * p, 0., p, -p};
        for (double d : edge) { ++n; if (tc::wrap_period(d, p) != ref_wrap(d, p)) ++bad; }
    }
    std::printf("in-box mismatches: %ld of %ld\n", bad, n);
    // far outside the box and periods far below ulp(d): must end, near [-p/2, p/2]
    long far_bad = 0;
    const double far[][2] = {{1e6, 1.}, {-7.3e9, 3.}, {1., 1e-20}, {-12345.678, 1e-15},
                             {1e300, 1e-300}, {5.5, 1.}, {-2.5000001, 1.}};
    for (auto& q : far) {
        const double r = tc::wrap_period(q[0], q[1]);
        // fmod reduces exactly, so the result lies in [-p/2, p/2]
        if (!(std::fabs(r) <= q[1] / 2.)) { ++far_bad; std::printf("far %g %g -> %g\n", q[0], q[1], r); }
    }
    std::printf("far mismatches: %ld\n", far_bad);
    return bad == 0 && far_bad == 0 ? 0 : 1;
}
