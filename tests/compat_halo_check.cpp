// Host check of the compat halos and neighbour table (one rank): after
// fill_halos every site of the extended (L+2)^4 box, faces, edges and corners
// included, holds the links of its periodic image in the interior -- what the
// reference's exchange_halos_cascade yields on a periodic lattice
// (src/mpi/HalosExchange.cpp:4-72) -- and the neighbour table
// (src/geometry/GeometryCB.cpp:10-31) reaches, in one and two hops, the sites
// whose links are those of x +- mu and of the staple's x + mu - nu and the
// clover's x - mu - nu.  Built with g++ against the compat sources by
// test_compat_halos.py.
#include <cstdio>

#include "gauge/GaugeField.h"

static Complex tag(int x, int y, int z, int t, int mu, int k) {
    return Complex(((t * 16 + z) * 16 + y) * 16 + x, mu * 9 + k);
}

int main() {
    const int L = 4;
    GeometryCB geo(L);
    GaugeField f(geo);
    for (int t = 1; t <= L; t++)
        for (int z = 1; z <= L; z++)
            for (int y = 1; y <= L; y++)
                for (int x = 1; x <= L; x++)
                    for (int mu = 0; mu < 4; mu++)
                        for (int k = 0; k < 9; k++)
                            f.view_link(geo.index(x, y, z, t), mu).data()[k] = tag(x, y, z, t, mu, k);
    f.fill_halos(geo);
    auto wrap = [&](int c) { return ((c - 1) % L + L) % L + 1; };  // any coordinate -> 1..L
    auto same_as = [&](size_t site, int x, int y, int z, int t) {
        int bad = 0;
        for (int mu = 0; mu < 4; mu++)
            for (int k = 0; k < 9; k++)
                bad += f.view_link_const(site, mu).data()[k] !=
                       tag(wrap(x), wrap(y), wrap(z), wrap(t), mu, k);
        return bad;
    };
    // 1. every site of the extended box is its periodic image
    int bad_box = 0;
    for (int t = 0; t <= L + 1; t++)
        for (int z = 0; z <= L + 1; z++)
            for (int y = 0; y <= L + 1; y++)
                for (int x = 0; x <= L + 1; x++) bad_box += same_as(geo.index(x, y, z, t), x, y, z, t);
    // 2. one and two hops through the neighbour table from every interior site
    int bad_hop = 0, checked = 0;
    for (int t = 1; t <= L; t++)
        for (int z = 1; z <= L; z++)
            for (int y = 1; y <= L; y++)
                for (int x = 1; x <= L; x++) {
                    const int c[4] = {x, y, z, t};
                    const size_t s = geo.index(x, y, z, t);
                    for (int mu = 0; mu < 4; mu++)
                        for (int d = 0; d < 2; d++) {
                            int e[4] = {c[0], c[1], c[2], c[3]};
                            e[mu] += d == up ? 1 : -1;
                            const size_t n1 = geo.get_neigh(s, mu, d == up ? up : down);
                            bad_hop += n1 != geo.index(e[0], e[1], e[2], e[3]);
                            bad_hop += same_as(n1, e[0], e[1], e[2], e[3]);
                            checked++;
                            // x + mu - nu (upward first) and x - mu - nu: the
                            // staple's and the clover's diagonal neighbours
                            for (int nu = 0; nu < 4; nu++) {
                                if (nu == mu) continue;
                                int g[4] = {e[0], e[1], e[2], e[3]};
                                g[nu] -= 1;
                                const size_t n2 = geo.get_neigh(n1, nu, down);
                                bad_hop += same_as(n2, g[0], g[1], g[2], g[3]);
                                checked++;
                            }
                        }
                }
    std::printf("halo sites differing from their periodic image: %d\n", bad_box);
    std::printf("neighbour hops checked %d, wrong: %d\n", checked, bad_hop);
    return bad_box == 0 && bad_hop == 0 ? 0 : 1;
}
