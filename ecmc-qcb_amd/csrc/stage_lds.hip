// stage_lds.hip -- RK3 stage kernel with the links that the four direction
// wavefronts of a tile share staged in LDS (FP64 and FP32-mixed).
//
// The plain stage kernel (flow.hip) has each direction wavefront of a
// 64-site tile load the 19 links around its link (16 with the x-row
// shuffles).  Several of those are the same link for different directions:
//   U_nu(x)        is read by all four wavefronts (own link, upper staples);
//   U_nu(x - nu)   is read by the three wavefronts mu != nu (lower staples);
// and when the wavefront holds whole x-rows, the x-neighbours of those are in
// the tile too.  On MI355X the vector-memory path (texture address and data
// units) of that kernel is about three quarters busy, so here the workgroup
// first copies these links into LDS with direct global->LDS loads (no
// registers held while they are in flight), meets at one barrier, and each
// wavefront then reads them, and their x-neighbours, from LDS:
//   own[nu] = rows 0-1 of U_nu(x),  dn[nu] = rows 0-1 of U_nu(x - nu),
// 2 x 4 x 6 KiB = 48 KiB per workgroup (three workgroups per CU).  The
// products are the same calls on the same words in the same order as the
// plain kernel's staple sum (GaugeField::compute_staple,
// src/gauge/GaugeField.cpp:74-97), so the flow is bit-identical to it.
//
// Needs full 64-site tiles of consecutive sites (V a multiple of 64) and one
// thread per link; the plain kernel runs otherwise.
#include "kernels.hpp"
#include "stage_lds.hpp"

namespace qcb {

template <prec P, int STAGE, bool XROWS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kLdsStageWaves<P>))) void
k_rk3_stage_lds(const typename storage<P>::vec* __restrict__ Uin,
                typename storage<P>::vec* __restrict__ Uout,
                typename storage<P>::avec* __restrict__ acc, lattice g, double cz, double ca,
                int* __restrict__ warn) {
    __shared__ lds_link<P> own[4], dn[4];
    const unsigned ntiles = g.V / 64u;
    const unsigned tile =
        tile_of_block(blockIdx.x, ntiles, g.map_mode, g.map_win, stage_tiles_per_t(g));
    if (tile >= ntiles) return;  // the whole workgroup: no barrier is skipped
    const unsigned lane = threadIdx.x;
    const int mu = (int)threadIdx.y;  // wavefront-uniform
    unsigned c[4];
    const unsigned s = tile_site(g, tile, lane, c);
    nbr n = neighbours_at(g, c);
    if constexpr (XROWS) set_x_lanes(n, g, lane);
    stage_tile_to_lds<P, XROWS>(Uin, g.V, mu, s, n, own, dn);
    __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0): the LDS copies have landed
    __syncthreads();
    const auto f = storage<P>::view(Uin, g.V);
    bool w = false;
    switch (mu) {
        case 0: stage_lds_body<P, STAGE, 0, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
        case 1: stage_lds_body<P, STAGE, 1, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
        case 2: stage_lds_body<P, STAGE, 2, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
        default: stage_lds_body<P, STAGE, 3, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
    }
    if (w) atomicOr(warn, 1);
}

// Both forms, with or without whole x-rows: without them the FP32-mixed
// stage at 48^4 runs in 1.17 ms against 1.38 ms for the plain kernel, FP64
// in 2.42 against 2.57 ms.
bool stage_lds_applies(prec p, const lattice& g) {
    (void)p;
    return g.lg_tile[0] < 0 && g.dpw == 1 && g.V % 64u == 0;
}

template <prec P>
static hipError_t stage_lds_launch(int stage, const void* Uin, void* Uout, void* acc,
                                   const lattice& g, double cz, double ca, int* warn,
                                   hipStream_t st) {
    using S = storage<P>;
    const auto* in = static_cast<const typename S::vec*>(Uin);
    auto* out = static_cast<typename S::vec*>(Uout);
    auto* a = static_cast<typename S::avec*>(acc);
    const unsigned ntiles = g.V / 64u;
    const dim3 block(kSitesPerBlock, 4);
    const dim3 grid(grid_for_tiles(ntiles, g.map_mode, g.map_win, stage_tiles_per_t(g)));
    const bool xr = x_rows_in_wave(g);
    switch (stage * 2 + (xr ? 1 : 0)) {
        case 2: k_rk3_stage_lds<P, 1, false><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 3: k_rk3_stage_lds<P, 1, true><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 4: k_rk3_stage_lds<P, 2, false><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 5: k_rk3_stage_lds<P, 2, true><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 6: k_rk3_stage_lds<P, 3, false><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 7: k_rk3_stage_lds<P, 3, true><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rk3_stage_lds(prec p, int stage, const void* Uin, void* Uout, void* acc,
                                const lattice& g, double cz, double ca, int* warn,
                                hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (!stage_lds_applies(p, g)) return hipErrorInvalidValue;
    return p == prec::f64 ? stage_lds_launch<prec::f64>(stage, Uin, Uout, acc, g, cz, ca, warn, st)
                          : stage_lds_launch<prec::f32m>(stage, Uin, Uout, acc, g, cz, ca, warn, st);
}

}  // namespace qcb
