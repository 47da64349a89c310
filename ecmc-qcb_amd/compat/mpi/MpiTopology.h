// MpiTopology.h -- drop-in for src/mpi/MpiTopology.h:16-27 on one GPU.
//
// The whole lattice lives on one MI355X, so the process is a single rank
// (rank 0 of size 1, even parity).  A caller that asks for n_core_dim = n > 1
// -- the reference's n^4 sub-domains of L_core^4 sites, as in its input files
// -- gets them as one lattice of n L_core sites per dimension: a GeometryCB
// built after this topology multiplies its extent by n (one_rank_fold), and a
// note says so on stderr.  The type stays because the reference's entry
// points take it.
#pragma once
#include <cstdio>
#include <stdexcept>

enum parity { even, odd };

namespace mpi {
// n_core_dim of the latest MpiTopology of this process (1 before any): the
// factor GeometryCB applies to its extent.
inline int& one_rank_fold() {
    static int fold = 1;
    return fold;
}

class MpiTopology {
public:
    int n_core_dim;
    int rank = 0, size = 1;
    int local_rank = 0;
    int local_coords[4] = {0, 0, 0, 0};
    parity p = even;
    int x0 = 0, xL = 0, y0 = 0, yL = 0, z0 = 0, zL = 0, t0 = 0, tL = 0;
    int cart_comm = 0;

    explicit MpiTopology(int n_core_dim_) : n_core_dim(n_core_dim_) {
        if (n_core_dim_ < 1) throw std::invalid_argument("MpiTopology: n_core_dim must be >= 1");
        one_rank_fold() = n_core_dim_;
        if (n_core_dim_ > 1)
            std::fprintf(stderr,
                         "MpiTopology: n_core_dim = %d on one GPU: the %d^4 sub-domains run as "
                         "one lattice of %d x L_core sites per dimension\n",
                         n_core_dim_, n_core_dim_, n_core_dim_);
    }
};
}  // namespace mpi
