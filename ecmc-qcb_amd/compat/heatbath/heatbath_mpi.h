// heatbath_mpi.h -- drop-in for src/heatbath/heatbath_mpi.h:13-17 on one GPU.
//
// samples() keeps the reference's signature and runs the Cabibbo-Marinari
// heatbath sweeps on the device (qcbf_heatbath): the host GaugeField is
// uploaded, swept N_samples * N_sweeps times and downloaded with periodic
// halos.  The whole lattice is one rank, whose parity is topo.p (even), so
// only calls with active_parity == topo.p update it, as in the reference's
// rank checkerboard.  The device draws its random numbers from a counter
// generator; each call keys it with one 64-bit draw of rng[0], so the host
// streams advance and a run is reproducible from the caller's seeds.  N_hits
// repeats the exact SU(2) draw on a link whose staple is fixed, which leaves
// the link's distribution unchanged; one draw per link and sweep is made.
// The sweep is the device checkerboard over (direction, site parity), so the
// extent must be even.  The per-link hit() and the host sweep() of the
// reference are not provided.
#pragma once
#include <random>
#include <vector>

#include "../gauge/GaugeField.h"
#include "../io/params.h"
#include "../mpi/MpiTopology.h"

namespace mpi::heatbathcb {
std::vector<double> samples(GaugeField& field, const GeometryCB& geo, MpiTopology& topo,
                            const HbParams& params, std::vector<std::mt19937_64>& rng,
                            parity active_parity);
}  // namespace mpi::heatbathcb
