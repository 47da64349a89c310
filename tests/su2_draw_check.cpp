// Host build of the device heatbath's SU(2) draw (su2_draw.hpp): writes n
// draws at one alpha with the method forced to stdout as raw doubles (four
// per draw) and the number of failed draws to stderr.  A checker for
// tests/test_su2_draw.py only.
//
// usage: su2_draw_check <alpha> <method: kp | creutz> <n> <seed>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "su2_draw.hpp"

int main(int argc, char* argv[]) {
    if (argc != 5) return 2;
    const double alpha = std::atof(argv[1]);
    const bool kp = std::strcmp(argv[2], "kp") == 0;
    const long n = std::atol(argv[3]);
    qcb::hb_rng r{qcb::hb_mix(std::strtoull(argv[4], nullptr, 10)), 0u};
    long failed = 0;
    for (long i = 0; i < n; i++) {
        double a[4];
        // creutz_below = 0 forces Kennedy-Pendleton, = +inf forces Creutz
        if (!qcb::su2_draw(alpha, r, a, kp ? 0.0 : 1e300)) {
            failed++;
            continue;
        }
        std::fwrite(a, sizeof(double), 4, stdout);
    }
    std::fprintf(stderr, "failed %ld\n", failed);
    return 0;
}
