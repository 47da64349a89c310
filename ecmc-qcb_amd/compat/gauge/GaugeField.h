// GaugeField.h -- drop-in for src/gauge/GaugeField.h:18-62 without Eigen.
//
// Same storage as the reference: std::vector<std::complex<double>> over the
// extended (L+2)^4 box, links[(site*4 + mu)*9 + 3i + j], identity at
// construction (GaugeField.h:30-43).  view_link returns a small row-major 3x3
// view with operator()(i, j) instead of an Eigen::Map; arithmetic on links
// belongs to the device library, so no matrix algebra is offered here.
#pragma once
#include <complex>
#include <cstddef>
#include <random>
#include <vector>

#include "../geometry/GeometryCB.h"

using Complex = std::complex<double>;

template <typename P>
struct LinkView3 {
    P* p;
    auto& operator()(int i, int j) const { return p[3 * i + j]; }
    P* data() const { return p; }
};
using LinkView = LinkView3<Complex>;
using ConstLinkView = LinkView3<const Complex>;

class GaugeField {
    int L_ext;
    int L_int;
    size_t V_ext;
    size_t V_int;

public:
    std::vector<Complex> links;

    explicit GaugeField(const GeometryCB& geo);

    LinkView view_link(size_t site, int mu) { return LinkView{&links[(site * 4 + mu) * 9]}; }
    [[nodiscard]] ConstLinkView view_link_const(size_t site, int mu) const {
        return ConstLinkView{&links[(site * 4 + mu) * 9]};
    }

    // Interior links in the reference's flat order (site x-fastest over
    // x,y,z,t = 1..L, mu, 3x3 row-major, re/im): V_int*72 doubles.
    void pack_interior(const GeometryCB& geo, double* out) const;
    void unpack_interior(const GeometryCB& geo, const double* in);
    // Periodic halos from the interior (what exchange_halos_cascade yields on
    // a single rank).
    void fill_halos(const GeometryCB& geo);
    // Haar-random links (GaugeField::hot_start, src/gauge/GaugeField.cpp:10-20),
    // generated on the device from a counter generator keyed by one draw of
    // rng; halos filled.
    void hot_start(const GeometryCB& geo, std::mt19937_64& rng);
};
