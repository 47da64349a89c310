#include "GeometryCB.h"

#include "../mpi/MpiTopology.h"

// Neighbour table over the extended (L+2)^4 box; links that would leave the
// box are SIZE_MAX, as in the reference (GeometryCB.cpp:10-31).  L_ is the
// caller's L_core; on one GPU the box holds all n_core_dim^4 sub-domains
// (mpi::one_rank_fold).
GeometryCB::GeometryCB(int L_) {
    L_int = L_ * mpi::one_rank_fold();
    L_ext = L_int + 2;
    V_int = static_cast<size_t>(L_int) * L_int * L_int * L_int;
    V_ext = static_cast<size_t>(L_ext) * L_ext * L_ext * L_ext;
    neighbors.assign(V_ext * 8, SIZE_MAX);
    parity.resize(V_ext);
    for (int t = 0; t < L_ext; t++)
        for (int z = 0; z < L_ext; z++)
            for (int y = 0; y < L_ext; y++)
                for (int x = 0; x < L_ext; x++) {
                    const int c[4] = {x, y, z, t};
                    const size_t s = index(x, y, z, t);
                    for (int mu = 0; mu < 4; mu++) {
                        int u[4] = {c[0], c[1], c[2], c[3]};
                        int d[4] = {c[0], c[1], c[2], c[3]};
                        u[mu] += 1;
                        d[mu] -= 1;
                        if (u[mu] < L_ext) neighbors[index_neigh(s, mu, up)] = index(u[0], u[1], u[2], u[3]);
                        if (d[mu] >= 0) neighbors[index_neigh(s, mu, down)] = index(d[0], d[1], d[2], d[3]);
                    }
                    parity[s] = ((x + y + z + t) % 2 == 0) ? EV : OD;
                }
}
