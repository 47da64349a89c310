// observables_mpi.h -- drop-in for src/observables/observables_mpi.h:8-19.
//
// The per-site helpers of the reference (clover_site, local_q_e_clover,
// levi_civita) are device code here and have no host entry point; the
// lattice-wide functions keep their signatures and conventions.
#pragma once
#include <utility>
#include <vector>

#include "../flow/gradient_flow.h"
#include "../gauge/GaugeField.h"
#include "../mpi/MpiTopology.h"

namespace mpi::observables {
// sum over interior sites and planes mu<nu of Re tr U_p / 3 (observables_mpi.cpp:11-35)
double mean_plaquette_local(const GaugeField& field, const GeometryCB& geo);
// mean plaquette (observables_mpi.cpp:39-48)
double mean_plaquette_global(GaugeField& field, const GeometryCB& geo, MpiTopology& topo);
// {sum_x q(x), sum_x e(x)} (observables_mpi.cpp:115-133)
std::pair<double, double> topo_q_e_clover(const GaugeField& field, const GeometryCB& geo);
// {Q, E} = {sum q, sum e / V} (observables_mpi.cpp:136-156)
std::pair<double, double> topo_q_e_clover_global(const GaugeField& field, const GeometryCB& geo,
                                                 MpiTopology& topo);
// flowed (t, Q, E) series (observables_mpi.cpp:159-186)
std::vector<double> topo_charge_flowed(GaugeField& field, const GeometryCB& geo, GradientFlow& gf,
                                       mpi::MpiTopology& topo, int N_steps_gf, int N_rk_steps);
}  // namespace mpi::observables
