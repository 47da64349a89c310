// Host-side check of the site-tile geometry and the workgroup -> tile orders
// of lattice.hpp: every stage tile shape covers each site exactly once with
// consistent coordinates, and every order (modes 0-3 over several windows and
// slab widths) visits each tile exactly once, padding slots aside.  Built
// with hipcc as host code and run by test_tile_maps.py.
#include <cstdio>
#include <vector>

#include "lattice.hpp"
using namespace qcb;

static int check_order(unsigned nt, unsigned per_t, int mode, unsigned win) {
    std::vector<int> seen(nt, 0);
    const unsigned G = grid_for_tiles(nt, mode, win, per_t);
    for (unsigned b = 0; b < G; b++) {
        const unsigned t = tile_of_block(b, nt, mode, win, per_t);
        if (t < nt) seen[t]++;
    }
    int bad = 0;
    for (unsigned t = 0; t < nt; t++) bad += seen[t] != 1;
    if (bad) std::printf("order mode %d win %u per_t %u ntiles %u: %d tiles missed or repeated\n",
                         mode, win, per_t, nt, bad);
    return bad;
}

// Stage workgroups with dpw directions per wavefront take a group of dpw
// tiles: over the grid and the four wavefronts of each workgroup, every tile
// must be reached by exactly 4/dpw wavefronts (one per block of directions).
static int check_groups(unsigned nt, unsigned per_t, int mode, unsigned win, unsigned dpw) {
    std::vector<int> seen(nt, 0);
    const tile_groups tg = group_tiles(nt, per_t, mode, win, dpw);
    const unsigned G = grid_for_tiles(tg.n, mode, tg.win, tg.per_t);
    for (unsigned b = 0; b < G; b++)
        for (unsigned y = 0; y < 4; y++) {
            const unsigned t = tile_in_group(tile_of_block(b, tg.n, mode, tg.win, tg.per_t), y, dpw);
            if (t < nt) seen[t]++;
        }
    int bad = 0;
    for (unsigned t = 0; t < nt; t++) bad += seen[t] != (int)(4 / dpw);
    if (bad)
        std::printf("groups of %u, mode %d win %u per_t %u ntiles %u: %d tiles missed or repeated\n",
                    dpw, mode, win, per_t, nt, bad);
    return bad;
}

// The default tile order of make_lattice: slab sweeps of at most 256 tiles
// where the t-slice allows it, the known choices at the benchmarked sizes.
static int check_default(int L, int mode, unsigned win) {
    const int d[4] = {L, L, L, L};
    const lattice g = make_lattice(d);
    const unsigned pt = site_tiles_per_t(g);
    int bad = g.map_mode != mode || g.map_win != win;
    if (g.map_mode == 3 && pt > 16u * 256u) bad += g.map_win > 256u;
    if (bad)
        std::printf("default order %d^4: mode %d win %u, expected %d %u\n", L, g.map_mode,
                    g.map_win, mode, win);
    return bad;
}

int main() {
    const int dims[][4] = {{32, 32, 32, 32}, {48, 48, 48, 48}, {16, 8, 8, 12}, {8, 8, 8, 8},
                           {12, 12, 12, 12}, {4, 6, 8, 10},     {6, 4, 5, 8},  {24, 24, 24, 24},
                           {64, 64, 64, 64}, {3, 5, 4, 7}};
    const int shapes[][4] = {{-1, 0, 0, 0}, {2, 2, 1, 1}, {3, 2, 1, 0}, {3, 3, 0, 0}, {4, 2, 0, 0},
                             {2, 2, 2, 0},  {1, 1, 2, 2}, {2, 1, 1, 2}, {5, 1, 0, 0}};
    const unsigned wins[] = {1, 3, 8, 27, 32, 64, 216, 256, 1024};
    int bad = 0, tried = 0;
    for (auto& d : dims) {
        for (auto& sh : shapes) {
            lattice g = make_lattice(d);
            const bool ok = sh[0] < 0 ? (set_tile_shape(g, nullptr), true) : set_tile_shape(g, sh);
            if (!ok) continue;
            tried++;
            std::vector<int> seen(g.V, 0);
            const unsigned nt = stage_tiles(g);
            for (unsigned t = 0; t < nt; t++)
                for (unsigned l = 0; l < 64; l++) {
                    unsigned c[4];
                    const unsigned s = tile_site(g, t, l, c);
                    if (s >= g.V) {
                        if (sh[0] >= 0) bad++;  // box tiles have no padding lanes
                        continue;
                    }
                    seen[s]++;
                    for (int k = 0; k < 4; k++) bad += c[k] >= (unsigned)g.L[k];
                    bad += ((c[3] * g.L[2] + c[2]) * g.L[1] + c[1]) * g.L[0] + c[0] != s;
                }
            for (unsigned s = 0; s < g.V; s++) bad += seen[s] != 1;
            const unsigned pt = stage_tiles_per_t(g);
            if (pt) bad += nt % pt != 0;
            for (int mode = 0; mode <= 3; mode++)
                for (unsigned w : wins) {
                    bad += check_order(nt, pt, mode, w);
                    for (unsigned dpw : {1u, 2u, 4u}) bad += check_groups(nt, pt, mode, w, dpw);
                }
        }
        // measurement tiles: 64 consecutive sites
        lattice g = make_lattice(d);
        const unsigned nm = (g.V + 63) / 64, pt = site_tiles_per_t(g);
        for (int mode = 0; mode <= 3; mode++)
            for (unsigned w : wins) bad += check_order(nm, pt, mode, w);
    }
    // 8^4 .. 64^4 keep 16 (or 8) slabs; 96^4, 104^4 and 112^4 get more, smaller ones
    bad += check_default(8, 3, 1) + check_default(12, 2, 1024) + check_default(16, 3, 4) +
           check_default(24, 3, 27) + check_default(32, 3, 32) + check_default(48, 3, 108) +
           check_default(64, 3, 256) + check_default(96, 3, 216) + check_default(104, 3, 138) +
           check_default(112, 3, 172);
    std::printf("checked %d lattice/shape pairs, %d errors\n", tried, bad);
    return bad != 0;
}
