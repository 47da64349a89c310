// A caller of the drop-in's per-site observables only (no GradientFlow):
// reads an ILDG configuration, prints clover_site for a few sites and planes
// (one on a halo site, one with mu == nu), local_q_e_clover, and checks
// levi_civita against the permutation sign; then returns from main without
// releasing anything (the drop-in's device lattice outlives static
// destruction safely).
//
// usage: obs_check <dirpath> <filename> <L>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gauge/GaugeField.h"
#include "io/ildg.h"
#include "mpi/MpiTopology.h"
#include "observables/observables_mpi.h"

int main(int argc, char* argv[]) {
    if (argc != 4) return 2;
    const int L = std::atoi(argv[3]);
    mpi::MpiTopology topo(1);
    GeometryCB geo(L);
    GaugeField field(geo);
    read_ildg_clime(argv[2], argv[1], field, geo, topo);
    const int pts[4][4] = {{1, 1, 1, 1}, {2, 3, 4, 5}, {L, 1, 2, L}, {0, 2, 2, 3}};  // last: halo
    const int planes[3][2] = {{0, 1}, {2, 3}, {1, 1}};
    for (const auto& p : pts) {
        const size_t s = geo.index(p[0], p[1], p[2], p[3]);
        for (const auto& pl : planes) {
            const SU3 F = mpi::observables::clover_site(field, geo, s, pl[0], pl[1]);
            std::printf("F %d %d %d %d %d %d", p[0], p[1], p[2], p[3], pl[0], pl[1]);
            for (int k = 0; k < 9; k++) std::printf(" %.17g %.17g", F.data()[k].real(), F.data()[k].imag());
            std::printf("\n");
        }
        const auto qe = mpi::observables::local_q_e_clover(field, geo, s);
        std::printf("QE %d %d %d %d %.17g %.17g\n", p[0], p[1], p[2], p[3], qe.first, qe.second);
    }
    bool ok = true;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            for (int c = 0; c < 4; c++)
                for (int d = 0; d < 4; d++) {
                    const int p[4] = {a, b, c, d};
                    int sign = 1;
                    for (int i = 0; i < 4; i++)
                        for (int j = i + 1; j < 4; j++) {
                            if (p[i] == p[j]) sign = 0;
                            else if (p[i] > p[j]) sign = -sign;
                        }
                    ok = ok && mpi::observables::levi_civita(a, b, c, d) == sign;
                }
    ok = ok && mpi::observables::levi_civita(0, 1, 2, 3) == 1 &&
         mpi::observables::levi_civita(1, 0, 2, 3) == -1 && mpi::observables::levi_civita(3, 2, 1, 0) == 1;
    std::printf("EPS %s\n", ok ? "ok" : "wrong");
    return 0;
}
