// batch.hpp -- ensembles of lattices of uneven volumes flowed together.
//
// Every configuration keeps its own structure-of-arrays field (fields.hpp)
// inside one allocation; a table of descriptors gives each its extents and
// offsets, and a flat list of 64-site tiles covers all of them.  One persistent
// kernel per RK3 stage walks that list, so an ensemble of 64 small lattices is
// three launches per step instead of 192, and small lattices no longer leave
// most CUs idle.
#pragma once
#include <hip/hip_runtime.h>

#include "fields.hpp"

namespace qcb {

struct cfg_desc {
    lattice g;          // extents, volume, strides of this configuration
    unsigned tile0;     // first tile of this configuration in the flat tile list
    unsigned ntiles;    // ceil(V / 64)
    size_t u_off;       // offset of its link field, in link-field words
    size_t a_off;       // offset of its su(3) accumulator, in su(3)-field words
};

// Launch shape of the persistent kernels.
struct batch_launch {
    unsigned grid;      // workgroups of the persistent grid; 0: one workgroup per tile
    int map_mode;       // tile order (tile_of_block)
    unsigned map_win;
    int dpw;            // directions per wavefront of the stage kernels (1, 2, 4)
    // stage kernels with each tile's shared links staged in LDS (flat grid,
    // dpw 1, every member a whole number of 64-consecutive-site tiles)
    int lds;
};

hipError_t launch_batch_stage(prec p, int stage, const void* Uin, void* Uout, void* acc,
                              const cfg_desc* cfgs, const unsigned* tile_cfg, unsigned ntiles,
                              const batch_launch& bl, double cz, double ca, int* warn,
                              hipStream_t st);

// Per-tile partial triples (plaquette, E, Q-density) into partials[3*ntiles],
// then one workgroup per configuration reduces its tiles in order and writes
// {P, E, Q} to out + 3*(slot*ncfg + c), slot = *slot_ctr (or 0 if slot_ctr is
// null); when bump is set the slot counter is advanced afterwards.
hipError_t launch_batch_measure(prec p, const void* U, const cfg_desc* cfgs,
                                const unsigned* tile_cfg, unsigned ntiles, int ncfg,
                                const batch_launch& bl, double* partials, double* out,
                                unsigned* slot_ctr, bool bump, hipStream_t st);

// RK3 stage 1 of every member fused with the measurement of its input; the
// measurement output is that of launch_batch_measure.  Every member must use
// 64-consecutive-site stage tiles.
hipError_t launch_batch_stage1_measure(prec p, const void* Uin, void* Uout, void* acc,
                                       const cfg_desc* cfgs, const unsigned* tile_cfg,
                                       unsigned ntiles, int ncfg, const batch_launch& bl,
                                       double cz, int* warn, double* partials, double* out,
                                       unsigned* slot_ctr, bool bump, hipStream_t st);

// slot counter reset by a kernel (graph-capturable, no memset node)
hipError_t launch_set_counter(unsigned* ctr, unsigned value, hipStream_t st);

}  // namespace qcb
