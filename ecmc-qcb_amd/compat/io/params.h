// params.h -- the parameter structs of src/io/params.h:17-23,46-63 that the
// heatbath driver's sampling loop reads (defaults as there).
#pragma once
#include <string>

struct HbParams {
    double beta = 6.0;
    int N_samples = 10;
    int N_hits = 1;
    int N_sweeps = 1;
    int N_therm = 100;
};

struct RunParamsHbCB {
    int L_core = 6;
    int n_core_dims = 2;
    bool cold_start = true;
    int N_switch_eo = 3;  // even/odd switches per shift
    int N_shift = 2;      // shifts
    int seed = 123;
    HbParams hp{};        // heatbath parameters of each even/odd update
    int N_therm = 100;    // thermalisation shifts
    bool topo = true;
    int N_shift_plaquette = 2;  // plaquette every N_shift_plaquette shifts
    int N_shift_topo = 10;      // topological charge every N_shift_topo shifts
    int N_steps_gf = 10;
    int N_rk_steps = 40;
    std::string run_name = "c";
    std::string run_dir = "data";
    int save_each_shifts = 2;
};
