// flow_stages.cpp -- a caller of the GradientFlow members one by one, as the
// reference's rk3_step sequences them (src/flow/gradient_flow.cpp:124-135):
// compute_force(0), compute_w1, compute_force(1), compute_w2,
// replace_force_0, compute_force(1), compute_w3.  It reads an ILDG
// configuration, runs that sequence on one GradientFlow and rk3_step on
// another, prints the largest difference of the two flowed fields, checks
// that rk3_step left periodic halos, writes the host mirrors (force_0 of the
// input, field_c after the step) as raw interior doubles for comparison, and
// saves the flowed field with save_ildg_clime as outdir/flowed/flowed.  It
// also prints the input's local and global plaquette, Q and E.
//
// usage: flow_stages <dirpath> <filename> <L> <outdir>
// (the configuration is read from dirpath/filename/filename)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../flow/gradient_flow.h"
#include "../gauge/GaugeField.h"
#include "../io/ildg.h"
#include "../mpi/MpiTopology.h"
#include "../observables/observables_mpi.h"

static bool write_raw(const std::string& path, const std::vector<double>& v) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(v.data(), sizeof(double), v.size(), fp) == v.size();
    return std::fclose(fp) == 0 && ok;
}

int main(int argc, char* argv[]) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <dirpath> <filename> <L> <outdir>\n", argv[0]);
        return 2;
    }
    const std::string filedir = argv[1], filename = argv[2], outdir = argv[4];
    const int L = std::atoi(argv[3]);
    mpi::MpiTopology topo(1);
    GeometryCB geo(L);
    GaugeField field(geo);
    read_ildg_clime(filename, filedir, field, geo, topo);
    // the lattice-wide observables of the input, local and global forms
    const auto qe_local = mpi::observables::topo_q_e_clover(field, geo);
    const auto qe_global = mpi::observables::topo_q_e_clover_global(field, geo, topo);
    std::printf("P_local %.17g\nQE_local %.17g %.17g\nQE_global %.17g %.17g\n",
                mpi::observables::mean_plaquette_local(field, geo), qe_local.first,
                qe_local.second, qe_global.first, qe_global.second);

    // stage by stage, host mirrors refreshed after every call
    GradientFlow a(0.02, field, geo);
    a.copy(field);
    a.compute_force(0);
    std::vector<double> force0(geo.V_int * 72), stepwise(geo.V_int * 72), fused(geo.V_int * 72);
    a.force_0.pack_interior(geo, force0.data());
    a.compute_w1();
    a.compute_force(1);
    a.compute_w2();
    a.replace_force_0();
    a.compute_force(1);
    a.compute_w3();
    a.field_c.pack_interior(geo, stepwise.data());

    // the same step through rk3_step
    GradientFlow b(0.02, field, geo);
    b.copy(field);
    b.rk3_step(topo);
    b.field_c.pack_interior(geo, fused.data());

    double dmax = 0.0;
    for (size_t k = 0; k < fused.size(); k++) dmax = std::fmax(dmax, std::fabs(fused[k] - stepwise[k]));
    std::printf("max |rk3_step - stage sequence| = %.3e\n", dmax);

    // rk3_step refreshes the halos of field_c: every halo site is its periodic image
    auto wrap = [&](int c) { return ((c - 1) % L + L) % L + 1; };
    long bad = 0;
    for (int t = 0; t <= L + 1; t++)
        for (int z = 0; z <= L + 1; z++)
            for (int y = 0; y <= L + 1; y++)
                for (int x = 0; x <= L + 1; x++)
                    for (int mu = 0; mu < 4; mu++)
                        for (int k = 0; k < 9; k++)
                            bad += b.field_c.view_link_const(geo.index(x, y, z, t), mu).data()[k] !=
                                   b.field_c.view_link_const(geo.index(wrap(x), wrap(y), wrap(z), wrap(t)), mu)
                                       .data()[k];
    std::printf("halo entries differing from their periodic image: %ld\n", bad);

    if (!write_raw(outdir + "/force0.bin", force0) || !write_raw(outdir + "/rk3.bin", fused) ||
        !write_raw(outdir + "/stepwise.bin", stepwise)) {
        std::fprintf(stderr, "cannot write to %s\n", outdir.c_str());
        return 1;
    }
    save_ildg_clime("flowed", outdir, b.field_c, geo, topo);
    return bad == 0 ? 0 : 1;
}
