#include "heatbath_mpi.h"

#include "../device_scratch.h"
#include "../flow/gradient_flow.h"

// Generates heatbath samples (src/heatbath/heatbath_mpi.cpp:60-75): for each
// of N_samples, N_sweeps full sweeps, on the device.
std::vector<double> mpi::heatbathcb::samples(GaugeField& field, const GeometryCB& geo,
                                             MpiTopology& topo, const HbParams& params,
                                             std::vector<std::mt19937_64>& rng,
                                             parity active_parity) {
    std::vector<double> meas(params.N_samples > 0 ? params.N_samples : 0);
    const int n = params.N_samples * params.N_sweeps;
    if (topo.p != active_parity || n <= 0) return meas;
    const uint64_t key = rng.empty() ? 0u : rng[0]();
    qcbf_flow* h = qcbcompat::upload(field, geo);
    qcbf_compat_check(qcbf_heatbath(h, params.beta, n, key), h, "qcbf_heatbath");
    qcbcompat::download(field, geo);
    return meas;
}
