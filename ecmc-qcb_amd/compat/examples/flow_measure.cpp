// flow_measure.cpp -- a caller written against the reference's own API, in the
// shape of src/bin/test.cpp:12-50 and the measurement call of
// src/bin/gauge_hb_cb.cpp:184-197: read an ILDG configuration, print its
// plaquette, run topo_charge_flowed, print (t, P-free) tQE at full precision
// and append the reference's "t Q E" text file.
//
// usage: flow_measure <dirpath> <filename> <L> <N_steps_gf> <N_rk_steps>
// (the configuration is read from dirpath/filename/filename)
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../flow/gradient_flow.h"
#include "../gauge/GaugeField.h"
#include "../io/ildg.h"
#include "../io/io.h"
#include "../mpi/MpiTopology.h"
#include "../observables/observables_mpi.h"

int main(int argc, char* argv[]) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s <dirpath> <filename> <L> <N_steps_gf> <N_rk_steps>\n", argv[0]);
        return 2;
    }
    const std::string filedir = argv[1], filename = argv[2];
    const int L_core = std::atoi(argv[3]);
    const int N_steps_gf = std::atoi(argv[4]), N_rk_steps = std::atoi(argv[5]);

    mpi::MpiTopology topo(1);
    GeometryCB geo(L_core);
    GaugeField field(geo);
    GradientFlow flow(0.02, field, geo);
    flow.mirror_host = false;

    read_ildg_clime(filename, filedir, field, geo, topo);
    const double p = mpi::observables::mean_plaquette_global(field, geo, topo);
    std::printf("P %.17g\n", p);

    std::vector<double> tQE =
        mpi::observables::topo_charge_flowed(field, geo, flow, topo, N_steps_gf, N_rk_steps);
    for (size_t i = 0; i + 2 < tQE.size(); i += 3)
        std::printf("tQE %.17g %.17g %.17g\n", tQE[i], tQE[i + 1], tQE[i + 2]);
    io::save_topo(tQE, filename, filedir, 10);
    return 0;
}
