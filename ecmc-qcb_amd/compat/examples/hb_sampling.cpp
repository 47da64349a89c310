// hb_sampling.cpp -- the sampling loop of the reference's heatbath driver
// (generate_hb_cb, src/bin/gauge_hb_cb.cpp:20-247, the loops at :104-197)
// written against the drop-in: hot start, thermalisation shifts, then for
// every shift a random global shift and N_switch_eo even/odd heatbath
// updates, the plaquette every N_shift_plaquette shifts and the flowed
// topological charge every N_shift_topo shifts through ONE GradientFlow object
// on the evolving configuration.  Before each flow measurement the
// configuration is saved with save_ildg_clime as <run_dir>/<run_name>_m<k>, so
// that it can be flowed again elsewhere; the (t, Q, E) series is printed at
// full precision and appended to the reference's "t Q E" text file.
//
// usage: hb_sampling <run_dir> <run_name> <L_core> <beta> <seed> <N_therm> <N_shift>
//                    <N_switch_eo> <N_sweeps> <N_shift_plaquette> <N_shift_topo>
//                    <N_steps_gf> <N_rk_steps> [<n_core_dims>]
// (the lattice is n_core_dims L_core sites per dimension; n_core_dims
// defaults to 1)
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../flow/gradient_flow.h"
#include "../gauge/GaugeField.h"
#include "../heatbath/heatbath_mpi.h"
#include "../io/ildg.h"
#include "../io/io.h"
#include "../io/params.h"
#include "../mpi/HalosShift.h"
#include "../mpi/MpiTopology.h"
#include "../mpi/Shift.h"
#include "../observables/observables_mpi.h"

int main(int argc, char* argv[]) {
    if (argc != 14 && argc != 15) {
        std::fprintf(stderr,
                     "usage: %s <run_dir> <run_name> <L_core> <beta> <seed> <N_therm> <N_shift> "
                     "<N_switch_eo> <N_sweeps> <N_shift_plaquette> <N_shift_topo> <N_steps_gf> "
                     "<N_rk_steps> [<n_core_dims>]\n",
                     argv[0]);
        return 2;
    }
    RunParamsHbCB rp;
    rp.run_dir = argv[1];
    rp.run_name = argv[2];
    rp.L_core = std::atoi(argv[3]);
    rp.n_core_dims = argc == 15 ? std::atoi(argv[14]) : 1;
    rp.hp.beta = std::atof(argv[4]);
    rp.seed = std::atoi(argv[5]);
    rp.N_therm = std::atoi(argv[6]);
    rp.N_shift = std::atoi(argv[7]);
    rp.N_switch_eo = std::atoi(argv[8]);
    rp.hp.N_samples = 1;
    rp.hp.N_sweeps = std::atoi(argv[9]);
    rp.N_shift_plaquette = std::atoi(argv[10]);
    rp.N_shift_topo = std::atoi(argv[11]);
    rp.N_steps_gf = std::atoi(argv[12]);
    rp.N_rk_steps = std::atoi(argv[13]);
    rp.cold_start = false;

    mpi::MpiTopology topo(rp.n_core_dims);
    GeometryCB geo(rp.L_core);
    GaugeField field(geo);
    std::vector<std::mt19937_64> rng(1);
    rng[0].seed(rp.seed);
    if (!rp.cold_start) field.hot_start(geo, rng[0]);
    const HbParams hp = rp.hp;
    HalosShift halo_shift(geo);
    GradientFlow flow(0.02, field, geo);
    flow.mirror_host = false;

    auto update = [&](void) {
        mpi::shift::random_shift(field, geo, halo_shift, topo, rng[0]);
        for (int j = 0; j < rp.N_switch_eo; j++) {
            mpi::heatbathcb::samples(field, geo, topo, hp, rng, even);
            mpi::heatbathcb::samples(field, geo, topo, hp, rng, odd);
        }
    };
    for (int i = 0; i < rp.N_therm; i++) update();

    int k = 0;
    for (int i = 0; i < rp.N_shift; i++) {
        update();
        if (i % rp.N_shift_plaquette == 0)
            std::printf("P %d %.17g\n", i, mpi::observables::mean_plaquette_global(field, geo, topo));
        if (i % rp.N_shift_topo == 0) {
            const std::string name = rp.run_name + "_m" + std::to_string(k);
            save_ildg_clime(name, rp.run_dir, field, geo, topo);
            const std::vector<double> tQE = mpi::observables::topo_charge_flowed(
                field, geo, flow, topo, rp.N_steps_gf, rp.N_rk_steps);
            for (size_t n = 0; n + 2 < tQE.size(); n += 3)
                std::printf("tQE %d %.17g %.17g %.17g\n", k, tQE[n], tQE[n + 1], tQE[n + 2]);
            io::save_topo(tQE, rp.run_name, rp.run_dir, 10);
            k++;
        }
    }
    mpi::observables::release_device_scratch();
    return 0;
}
