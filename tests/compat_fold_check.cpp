// Host check of the one-GPU topology of the drop-in: MpiTopology(n) followed
// by GeometryCB(L_core) gives one lattice of n * L_core sites per dimension,
// MpiTopology(1) restores L_core, and n < 1 is rejected.
#include <cstdio>
#include <stdexcept>

#include "geometry/GeometryCB.h"
#include "mpi/MpiTopology.h"

int main() {
    int wrong = 0;
    {
        const GeometryCB g(4);
        wrong += g.L_int != 4;
    }
    {
        mpi::MpiTopology topo(2);
        const GeometryCB g(4);
        wrong += g.L_int != 8 || g.V_int != 4096 || g.L_ext != 10 || topo.size != 1 || topo.rank != 0;
        const size_t s = g.index(8, 8, 8, 8);  // last interior site
        wrong += g.get_neigh(s, 0, up) != g.index(9, 8, 8, 8);
    }
    {
        mpi::MpiTopology topo(1);
        const GeometryCB g(6);
        wrong += g.L_int != 6;
    }
    bool threw = false;
    try {
        mpi::MpiTopology bad(0);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    wrong += !threw;
    std::printf("wrong: %d\n", wrong);
    return wrong != 0;
}
