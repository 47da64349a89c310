#include "observables_mpi.h"

#include <iostream>
#include <memory>

#include "../device_scratch.h"
#include "../io/io.h"

namespace {
qcbf_flow* upload(const GaugeField& field, const GeometryCB& geo) {
    return qcbcompat::upload(field, geo);
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

int mpi::observables::levi_civita(int mu, int nu, int rho, int sigma) {
    const int p[4] = {mu, nu, rho, sigma};
    int sign = 1;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) {
            if (p[i] == p[j]) return 0;
            if (p[i] > p[j]) sign = -sign;
        }
    return sign;
}

SU3 mpi::observables::clover_site(const GaugeField& field, const GeometryCB& geo, size_t site,
                                  int mu, int nu) {
    if (mu == nu) std::cerr << "mu = nu => G(site, mu, nu) = 0\n";  // observables_mpi.cpp:68
    qcbf_flow* h = upload(field, geo);
    double out[18];
    qcbf_compat_check(qcbf_clover_site(h, qcbcompat::interior_site(geo, site), mu, nu, out), h,
                      "qcbf_clover_site");
    SU3 F;
    for (int k = 0; k < 9; k++) F.data()[k] = Complex(out[2 * k], out[2 * k + 1]);
    return F;
}

std::pair<double, double> mpi::observables::local_q_e_clover(const GaugeField& field,
                                                             const GeometryCB& geo, size_t site) {
    qcbf_flow* h = upload(field, geo);
    double q = 0.0, e = 0.0;
    qcbf_compat_check(qcbf_local_q_e(h, qcbcompat::interior_site(geo, site), &q, &e), h,
                      "qcbf_local_q_e");
    return {q, e};
}

void mpi::observables::release_device_scratch() { qcbcompat::release_device_scratch(); }

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
