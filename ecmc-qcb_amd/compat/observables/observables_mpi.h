// observables_mpi.h -- drop-in for src/observables/observables_mpi.h:8-19.
//
// Every function keeps the reference's signature and conventions and
// computes on the device: a host GaugeField is uploaded to the drop-in's
// device lattice (device_scratch.h) on each call, so the per-site functions
// cost one upload each; a caller that needs many sites of one field can call
// qcbf_clover_site / qcbf_local_q_e on a handle directly.
#pragma once
#include <utility>
#include <vector>

#include "../flow/gradient_flow.h"
#include "../gauge/GaugeField.h"
#include "../mpi/MpiTopology.h"
#include "../su3/utils.h"

namespace mpi::observables {
// sum over interior sites and planes mu<nu of Re tr U_p / 3 (observables_mpi.cpp:11-35)
double mean_plaquette_local(const GaugeField& field, const GeometryCB& geo);
// mean plaquette (observables_mpi.cpp:39-48)
double mean_plaquette_global(GaugeField& field, const GeometryCB& geo, MpiTopology& topo);
// F_munu(site) = (C - C^dag)/8 of the four-leaf clover (observables_mpi.cpp:66-90);
// site is an index of geo's extended box (a halo site means its periodic image)
SU3 clover_site(const GaugeField& field, const GeometryCB& geo, size_t site, int mu, int nu);
// sign of the permutation (mu, nu, rho, sigma) of (0, 1, 2, 3), 0 if two are
// equal (observables_mpi.cpp:51-63)
int levi_civita(int mu, int nu, int rho, int sigma);
// {q(site), e(site)} (observables_mpi.cpp:93-112)
std::pair<double, double> local_q_e_clover(const GaugeField& field, const GeometryCB& geo,
                                           size_t site);
// {sum_x q(x), sum_x e(x)} (observables_mpi.cpp:115-133)
std::pair<double, double> topo_q_e_clover(const GaugeField& field, const GeometryCB& geo);
// {Q, E} = {sum q, sum e / V} (observables_mpi.cpp:136-156)
std::pair<double, double> topo_q_e_clover_global(const GaugeField& field, const GeometryCB& geo,
                                                 MpiTopology& topo);
// flowed (t, Q, E) series (observables_mpi.cpp:159-186)
std::vector<double> topo_charge_flowed(GaugeField& field, const GeometryCB& geo, GradientFlow& gf,
                                       mpi::MpiTopology& topo, int N_steps_gf, int N_rk_steps);
// Frees the device lattice the functions above upload host fields to (it is
// otherwise kept until the process ends; no counterpart in the reference).
void release_device_scratch();
}  // namespace mpi::observables
