#include "Shift.h"

#include <algorithm>
#include <iostream>

#include "HalosExchange.h"

namespace {
// Visits the interior sites whose coordinate along dim lies in [lo, hi]
// (1-based), t outermost and x fastest, as f(site, c) with c the coordinates.
template <class F>
void for_slices(const GeometryCB& geo, int dim, int lo, int hi, F&& f) {
    int a[4] = {1, 1, 1, 1}, b[4] = {geo.L_int, geo.L_int, geo.L_int, geo.L_int};
    a[dim] = lo;
    b[dim] = hi;
    int c[4];
    for (c[3] = a[3]; c[3] <= b[3]; c[3]++)
        for (c[2] = a[2]; c[2] <= b[2]; c[2]++)
            for (c[1] = a[1]; c[1] <= b[1]; c[1]++)
                for (c[0] = a[0]; c[0] <= b[0]; c[0]++) f(geo.index(c[0], c[1], c[2], c[3]), c);
}
}  // namespace

void mpi::shift::pack_shift_slice(const GaugeField& field, const GeometryCB& geo,
                                  std::vector<Complex>& buffer, int dim, int start, int end) {
    size_t k = 0;
    for_slices(geo, dim, start, end, [&](size_t site, const int*) {
        const Complex* src = field.view_link_const(site, 0).data();
        std::copy(src, src + 36, buffer.begin() + k);
        k += 36;
    });
}

void mpi::shift::unpack_shift_slice(GaugeField& field, const GeometryCB& geo,
                                    const std::vector<Complex>& buffer, int dim, int start,
                                    int end) {
    size_t k = 0;
    for_slices(geo, dim, start, end, [&](size_t site, const int*) {
        std::copy(buffer.begin() + k, buffer.begin() + k + 36, field.view_link(site, 0).data());
        k += 36;
    });
}

void mpi::shift::apply_local_shift(GaugeField& field, const GeometryCB& geo, int dim, int s) {
    if (s <= 0) return;
    // highest slices first, so every source is read before it is overwritten
    for (int c = geo.L_int; c > s; c--)
        for_slices(geo, dim, c, c, [&](size_t dst, const int* co) {
            int src[4] = {co[0], co[1], co[2], co[3]};
            src[dim] -= s;
            const Complex* from = field.view_link_const(geo.index(src[0], src[1], src[2], src[3]), 0).data();
            std::copy(from, from + 36, field.view_link(dst, 0).data());
        });
}

void mpi::shift::shift_field(GaugeField& field, const GeometryCB& geo, HalosShift& h,
                             mpi::MpiTopology&, const int s[4]) {
    for (int dim = 0; dim < 4; dim++) {
        const int sh = s[dim];
        if (sh == 0) continue;
        pack_shift_slice(field, geo, h.send, dim, geo.L_int - sh + 1, geo.L_int);
        // one rank: the slices leaving at the top come back in at the bottom
        std::copy(h.send.begin(), h.send.begin() + static_cast<size_t>(sh) * geo.L_int * geo.L_int *
                                                      geo.L_int * 36,
                  h.recv.begin());
        apply_local_shift(field, geo, dim, sh);
        unpack_shift_slice(field, geo, h.recv, dim, 1, sh);
    }
}

void mpi::shift::random_shift(GaugeField& field, const GeometryCB& geo, HalosShift& h,
                              mpi::MpiTopology& topo, std::mt19937_64& rng) {
    std::uniform_int_distribution<int> pick(0, geo.L_int - 1);
    int s[4];
    for (int& v : s) v = pick(rng);  // x, y, z, t, as Shift.cpp:166-170
    if (topo.rank == 0)
        std::cout << "Shift (" << s[0] << ", " << s[1] << ", " << s[2] << ", " << s[3] << ")\n";
    shift_field(field, geo, h, topo, s);
    mpi::exchange::exchange_halos_cascade(field, geo, topo);
}
