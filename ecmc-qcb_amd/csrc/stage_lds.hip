// stage_lds.hip -- FP64 RK3 stage kernel with the links that the four
// direction wavefronts of a tile share staged in LDS.
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
// thread per link; the launcher falls back to the plain kernel otherwise.
#include "flow_body.hpp"
#include "kernels.hpp"

namespace qcb {

// Rows 0-1 of one link for the 64 sites of a tile: [word][lane].
struct lds_link {
    double2 w[6][64];
};

// Lane i copies rows 0-1 of U_mu(site_i) into L.w[.][i] (global_load_lds:
// lane i's 16 bytes land at the wave-uniform LDS base + 16 i).
__device__ inline void stage_to_lds(const double2* __restrict__ U, unsigned V, int mu,
                                    unsigned site, lds_link& L) {
    const double2* b = U + (size_t)mu * 9 * V + site;
#pragma unroll
    for (int k = 0; k < 6; k++)
        __builtin_amdgcn_global_load_lds((const void*)(b + (size_t)k * V),
                                         (__attribute__((address_space(3))) void*)&L.w[k][0], 16,
                                         0, 0);
}

__device__ inline mat3<double> from_lds(const lds_link& L, unsigned lane) {
    double2 w[6];
#pragma unroll
    for (int k = 0; k < 6; k++) w[k] = L.w[k][lane];
    return field64::from_words(w);
}

// Staple sum of U_MU(x) with the shared links from LDS.  XROWS: the tile is
// whole x-rows (n.xp, n.xm are the lanes of x+1, x-1), so the x-neighbours of
// the staged links are staged too.  Same operands and products, in the same
// order, as staple_sum (flow_body.hpp).
template <int MU, bool XROWS>
__device__ inline mat3<double> staple_sum_lds(const field64& f, unsigned s, const nbr& n,
                                              const lds_link* own, const lds_link* dn,
                                              unsigned lane) {
    mat3<double> st = zero3<double>();
    const unsigned xmu = s + n.up[MU];
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == MU) continue;
        const unsigned xnu = s + n.up[nu];
        const unsigned xmnu = s + n.dn[nu];
        const unsigned xmunu = xmu + n.dn[nu];
        {  // U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag
            const mat3<double> c = (XROWS && nu == 0) ? from_lds(own[MU], n.xp) : f.link(MU, xnu);
            const mat3<double> b = mul_su3(from_lds(own[nu], lane), c);
            const mat3<double> a = (XROWS && MU == 0) ? from_lds(own[nu], n.xp) : f.link(nu, xmu);
            add_to(st, mul_nd_su3(a, b));
        }
        {  // (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
            const mat3<double> d = (XROWS && nu == 0) ? from_lds(own[MU], n.xm) : f.link(MU, xmnu);
            const mat3<double> e = (XROWS && MU == 0) ? from_lds(dn[nu], n.xp) : f.link(nu, xmunu);
            const mat3<double> b = mul_su3(d, e);
            const mat3<double> v = (XROWS && nu == 0) ? from_lds(own[0], n.xm) : from_lds(dn[nu], lane);
            add_to(st, mul_dn_su3(b, v));
        }
        // keep the next plane's loads below this plane's arithmetic
        __builtin_amdgcn_sched_barrier(0);
    }
    return st;
}

template <int STAGE, int MU, bool XROWS>
__device__ inline void stage_lds_body(const field64& f, double2* __restrict__ Uout,
                                      double2* __restrict__ acc, unsigned s, const nbr& n,
                                      const lds_link* own, const lds_link* dn, unsigned lane,
                                      double cz, double ca, bool& warn) {
    const mat3<double> st = staple_sum_lds<MU, XROWS>(f, s, n, own, dn, lane);
    const mat3<double> u = from_lds(own[MU], lane);
    const alg8<double> z = neg_proj_ta(mul(u, st));
    stage_update<prec::f64, STAGE, MU>(z, u, f.V, Uout, acc, s, cz, ca, warn);
}

template <int STAGE, bool XROWS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_rk3_stage_lds(
    const double2* __restrict__ Uin, double2* __restrict__ Uout, double2* __restrict__ acc,
    lattice g, double cz, double ca, int* __restrict__ warn) {
    __shared__ lds_link own[4], dn[4];
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
    // Stage U_mu(x) and U_nu(x - nu).  With whole x-rows dn[0] is own[0] at
    // x-1 and is not staged; wavefront 0, which loads the fewest links later,
    // stages dn[1..3].  Otherwise wavefront mu stages dn[mu].
    stage_to_lds(Uin, g.V, mu, s, own[mu]);
    if constexpr (XROWS) {
        if (mu == 0)
#pragma unroll
            for (int nu = 1; nu < 4; nu++) stage_to_lds(Uin, g.V, nu, s + n.dn[nu], dn[nu]);
    } else {
        stage_to_lds(Uin, g.V, mu, s + n.dn[mu], dn[mu]);
    }
    __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0): the LDS copies have landed
    __syncthreads();
    const field64 f{Uin, g.V};
    bool w = false;
    switch (mu) {
        case 0: stage_lds_body<STAGE, 0, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
        case 1: stage_lds_body<STAGE, 1, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
        case 2: stage_lds_body<STAGE, 2, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
        default: stage_lds_body<STAGE, 3, XROWS>(f, Uout, acc, s, n, own, dn, lane, cz, ca, w); break;
    }
    if (w) atomicOr(warn, 1);
}

bool stage_lds_applies(const lattice& g) {
    return g.lg_tile[0] < 0 && g.dpw == 1 && g.V % 64u == 0;
}

hipError_t launch_rk3_stage_lds(int stage, const void* Uin, void* Uout, void* acc,
                                const lattice& g, double cz, double ca, int* warn,
                                hipStream_t st) {
    if (!stage_lds_applies(g)) return hipErrorInvalidValue;
    const auto* in = static_cast<const double2*>(Uin);
    auto* out = static_cast<double2*>(Uout);
    auto* a = static_cast<double2*>(acc);
    const unsigned ntiles = g.V / 64u;
    const dim3 block(kSitesPerBlock, 4);
    const dim3 grid(grid_for_tiles(ntiles, g.map_mode, g.map_win, stage_tiles_per_t(g)));
    const bool xr = x_rows_in_wave(g);
    switch (stage * 2 + (xr ? 1 : 0)) {
        case 2: k_rk3_stage_lds<1, false><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 3: k_rk3_stage_lds<1, true><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 4: k_rk3_stage_lds<2, false><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 5: k_rk3_stage_lds<2, true><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 6: k_rk3_stage_lds<3, false><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 7: k_rk3_stage_lds<3, true><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace qcb
