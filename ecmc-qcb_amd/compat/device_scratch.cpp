#include "device_scratch.h"

#include "flow/gradient_flow.h"

namespace {
struct scratch_lattice {
    int L = 0;
    qcbf_flow* h = nullptr;
    std::vector<double> buf;
};
// never destroyed: see device_scratch.h
scratch_lattice& state() {
    static scratch_lattice* s = new scratch_lattice;
    return *s;
}
}  // namespace

qcbf_flow* qcbcompat::scratch(const GeometryCB& geo) {
    scratch_lattice& s = state();
    if (s.h && s.L == geo.L_int) return s.h;
    release_device_scratch();
    const int dims[4] = {geo.L_int, geo.L_int, geo.L_int, geo.L_int};
    qcbf_compat_check(qcbf_create(dims, QCBF_FP64, &s.h), nullptr, "qcbf_create");
    s.L = geo.L_int;
    s.buf.resize(geo.V_int * 72);
    return s.h;
}

qcbf_flow* qcbcompat::upload(const GaugeField& field, const GeometryCB& geo) {
    qcbf_flow* h = scratch(geo);
    scratch_lattice& s = state();
    field.pack_interior(geo, s.buf.data());
    qcbf_compat_check(qcbf_set_links(h, s.buf.data()), h, "qcbf_set_links");
    return h;
}

void qcbcompat::download(GaugeField& field, const GeometryCB& geo) {
    qcbf_flow* h = scratch(geo);
    scratch_lattice& s = state();
    qcbf_compat_check(qcbf_get_links(h, s.buf.data()), h, "qcbf_get_links");
    field.unpack_interior(geo, s.buf.data());
    field.fill_halos(geo);
}

void qcbcompat::release_device_scratch() {
    scratch_lattice& s = state();
    if (s.h) qcbf_destroy(s.h);
    s.h = nullptr;
    s.L = 0;
    s.buf.clear();
    s.buf.shrink_to_fit();
}

long qcbcompat::interior_site(const GeometryCB& geo, size_t site) {
    const int Le = geo.L_ext, L = geo.L_int;
    int c[4];
    size_t r = site;
    for (int d = 0; d < 4; d++) {
        const int v = static_cast<int>(r % Le);
        r /= Le;
        c[d] = v == 0 ? L - 1 : (v == Le - 1 ? 0 : v - 1);  // halo -> periodic image
    }
    return ((static_cast<long>(c[3]) * L + c[2]) * L + c[1]) * L + c[0];
}

// GaugeField::hot_start lives here, beside the device lattice it uses, so that
// GaugeField.cpp stays host-only.
void GaugeField::hot_start(const GeometryCB& geo, std::mt19937_64& rng) {
    qcbf_flow* h = qcbcompat::scratch(geo);
    qcbf_compat_check(qcbf_hot_start(h, rng()), h, "qcbf_hot_start");
    qcbcompat::download(*this, geo);
}
