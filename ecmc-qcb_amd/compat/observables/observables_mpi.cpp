#include "observables_mpi.h"

#include <iostream>
#include <memory>

#include "../io/io.h"

namespace {
// A device lattice kept between calls for observables of host fields.
struct scratch_lattice {
    int L = 0;
    qcbf_flow* h = nullptr;
    std::vector<double> buf;
    ~scratch_lattice() { qcbf_destroy(h); }
};

qcbf_flow* upload(const GaugeField& field, const GeometryCB& geo) {
    static scratch_lattice s;
    if (s.L != geo.L_int) {
        qcbf_destroy(s.h);
        s.h = nullptr;
        const int dims[4] = {geo.L_int, geo.L_int, geo.L_int, geo.L_int};
        qcbf_compat_check(qcbf_create(dims, QCBF_FP64, &s.h), nullptr, "qcbf_create");
        s.L = geo.L_int;
        s.buf.resize(geo.V_int * 72);
    }
    field.pack_interior(geo, s.buf.data());
    qcbf_compat_check(qcbf_set_links(s.h, s.buf.data()), s.h, "qcbf_set_links");
    return s.h;
}
}  // namespace

double mpi::observables::mean_plaquette_local(const GaugeField& field, const GeometryCB& geo) {
    qcbf_flow* h = upload(field, geo);
    double p = 0.0;
    qcbf_compat_check(qcbf_plaquette(h, &p), h, "qcbf_plaquette");
    return p * 6.0 * static_cast<double>(geo.V_int);
}

double mpi::observables::mean_plaquette_global(GaugeField& field, const GeometryCB& geo,
                                               MpiTopology&) {
    qcbf_flow* h = upload(field, geo);
    double p = 0.0;
    qcbf_compat_check(qcbf_plaquette(h, &p), h, "qcbf_plaquette");
    return p;
}

std::pair<double, double> mpi::observables::topo_q_e_clover(const GaugeField& field,
                                                            const GeometryCB& geo) {
    qcbf_flow* h = upload(field, geo);
    double q = 0.0, e = 0.0;
    qcbf_compat_check(qcbf_q_e_clover(h, &q, &e), h, "qcbf_q_e_clover");
    return {q, e * static_cast<double>(geo.V_int)};
}

std::pair<double, double> mpi::observables::topo_q_e_clover_global(const GaugeField& field,
                                                                   const GeometryCB& geo,
                                                                   MpiTopology&) {
    qcbf_flow* h = upload(field, geo);
    double q = 0.0, e = 0.0;
    qcbf_compat_check(qcbf_q_e_clover(h, &q, &e), h, "qcbf_q_e_clover");
    return {q, e};
}

std::vector<double> mpi::observables::topo_charge_flowed(GaugeField& field, const GeometryCB& geo,
                                                         GradientFlow& gf, mpi::MpiTopology& topo,
                                                         int N_steps_gf, int N_rk_steps) {
    std::vector<double> tQE(3 * static_cast<size_t>(N_steps_gf > 0 ? N_steps_gf : 0));
    gf.copy(field);
    qcbf_compat_check(qcbf_flow_measure(gf.device(), gf.epsilon, N_steps_gf, N_rk_steps,
                                        tQE.data(), nullptr),
                      gf.device(), "qcbf_flow_measure");
    if (topo.rank == 0) {
        const int p = 6, pt = 2;
        for (int n = 0; n < N_steps_gf; n++) {
            const double t = tQE[3 * n], Q = tQE[3 * n + 1], E = tQE[3 * n + 2];
            std::cout << "n = " << n << ", t = " << io::format_double(t, pt)
                      << ", Q = " << io::format_double(Q, p)
                      << ", t²E = " << io::format_double(t * t * E, p) << "\n";
        }
    }
    gf.download_field_c();
    gf.field_c.fill_halos(geo);
    return tQE;
}
