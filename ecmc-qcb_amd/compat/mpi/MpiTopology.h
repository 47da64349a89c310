// MpiTopology.h -- drop-in for src/mpi/MpiTopology.h:16-27 on one GPU.
//
// The whole lattice lives on one MI355X, so the process grid is a single
// rank: n_core_dim must be 1 (the constructor throws std::invalid_argument
// otherwise).  The type stays because the reference's entry points take it.
#pragma once
#include <stdexcept>

enum parity { even, odd };

namespace mpi {
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
        if (n_core_dim_ != 1)
            throw std::invalid_argument(
                "MpiTopology: the GPU drop-in holds the whole lattice on one device; "
                "use n_core_dim = 1 and L_core = global extent");
    }
};
}  // namespace mpi
