// Host check of the compat mpi::shift (one rank): shift_field moves every link
// by s with periodic wrap, pack/unpack round-trip, and random_shift draws the
// caller's std::mt19937_64 stream as uniform ints in [0, L-1], x then y, z, t.
// Built with g++ against the compat sources by test_compat_shift.py.
#include <cstdio>
#include <random>

#include "gauge/GaugeField.h"
#include "mpi/Shift.h"

static Complex tag(int x, int y, int z, int t, int mu, int k) {
    return Complex(((t * 16 + z) * 16 + y) * 16 + x, mu * 9 + k);
}

int main() {
    const int L = 5;
    GeometryCB geo(L);
    GaugeField f(geo);
    auto fill = [&](GaugeField& g) {
        for (int t = 1; t <= L; t++)
            for (int z = 1; z <= L; z++)
                for (int y = 1; y <= L; y++)
                    for (int x = 1; x <= L; x++)
                        for (int mu = 0; mu < 4; mu++)
                            for (int k = 0; k < 9; k++)
                                g.view_link(geo.index(x, y, z, t), mu).data()[k] =
                                    tag(x, y, z, t, mu, k);
    };
    fill(f);
    mpi::MpiTopology topo(1);
    HalosShift h(geo);
    const int s[4] = {1, 3, 0, 4};
    mpi::shift::shift_field(f, geo, h, topo, s);
    int bad = 0;
    auto back = [&](int c, int d) { return ((c - 1 - s[d]) % L + L) % L + 1; };
    for (int t = 1; t <= L; t++)
        for (int z = 1; z <= L; z++)
            for (int y = 1; y <= L; y++)
                for (int x = 1; x <= L; x++)
                    for (int mu = 0; mu < 4; mu++)
                        for (int k = 0; k < 9; k++)
                            bad += f.view_link(geo.index(x, y, z, t), mu).data()[k] !=
                                   tag(back(x, 0), back(y, 1), back(z, 2), back(t, 3), mu, k);
    std::printf("shift_field mismatches: %d\n", bad);

    // random_shift: same draws as the caller's generator; halos refreshed
    GaugeField g(geo);
    fill(g);
    std::mt19937_64 rng(2024), mirror(2024);
    mpi::shift::random_shift(g, geo, h, topo, rng);
    std::uniform_int_distribution<int> pick(0, L - 1);
    int r[4];
    for (int& v : r) v = pick(mirror);
    GaugeField e(geo);
    fill(e);
    mpi::shift::shift_field(e, geo, h, topo, r);
    e.fill_halos(geo);
    int bad2 = 0;
    for (size_t i = 0; i < g.links.size(); i++) bad2 += g.links[i] != e.links[i];
    bad2 += rng() != mirror();
    std::printf("random_shift mismatches: %d\n", bad2);
    return bad + bad2 != 0;
}
