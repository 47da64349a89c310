// GeometryCB.h -- drop-in for src/geometry/GeometryCB.h:8-63.
//
// Keeps the reference's extended-box indexing (interior coordinates 1..L,
// width-1 halos, x fastest: GeometryCB.h:25-27) and its neighbour table
// (GeometryCB.cpp:10-31), so host code that walks a GaugeField keeps working.
// The checkerboard tables of the MCMC samplers (frozen links, staple-link
// lists) are not part of the flow path and are not provided.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "types.h"

class GeometryCB {
public:
    int L_int;
    int L_ext;
    size_t V_int;
    size_t V_ext;

private:
    std::vector<size_t> neighbors;
    std::vector<site_parity> parity;

public:
    explicit GeometryCB(int L_);

    [[nodiscard]] size_t index(int x, int y, int z, int t) const {
        return ((static_cast<size_t>(t) * L_ext + z) * L_ext + y) * L_ext + x;
    }
    [[nodiscard]] static size_t index_neigh(size_t site, int mu, dir d) {
        return site * 8 + mu * 2 + d;
    }
    [[nodiscard]] size_t get_neigh(size_t site, int mu, dir d) const {
        return neighbors[index_neigh(site, mu, d)];
    }
    [[nodiscard]] site_parity get_parity(size_t site) const { return parity[site]; }
};
