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
#include "flow_body.hpp"
#include "kernels.hpp"

namespace qcb {

// Rows 0-1 of one link for the 64 sites of a tile, as the field stores them
// (FP64: six double2, FP32-mixed: three float4 per link): [word][lane].
template <prec P>
struct lds_link {
    using F = typename storage<P>::field;
    typename F::word w[F::kWords][64];
};

// Word k of U_mu(site) in the field's structure of arrays (fields.hpp).
template <prec P>
__device__ inline const typename storage<P>::vec* link_word(const typename storage<P>::vec* U,
                                                            unsigned V, int mu, int k,
                                                            unsigned site) {
    constexpr int stride = P == prec::f64 ? 9 : 3;  // words per link in the field
    return U + ((size_t)mu * stride + k) * V + site;
}

// Lane i copies rows 0-1 of U_mu(site_i) into L.w[.][i] (global_load_lds:
// lane i's 16 bytes land at the wave-uniform LDS base + 16 i).
template <prec P>
__device__ inline void stage_to_lds(const typename storage<P>::vec* __restrict__ U, unsigned V,
                                    int mu, unsigned site, lds_link<P>& L) {
    static_assert(sizeof(typename storage<P>::vec) == 16, "16-byte link words");
#pragma unroll
    for (int k = 0; k < storage<P>::field::kWords; k++)
        __builtin_amdgcn_global_load_lds((const void*)link_word<P>(U, V, mu, k, site),
                                         (__attribute__((address_space(3))) void*)&L.w[k][0], 16,
                                         0, 0);
}

template <prec P>
__device__ inline mat3<typename storage<P>::real> from_lds(const lds_link<P>& L, unsigned lane) {
    using F = typename storage<P>::field;
    typename F::word w[F::kWords];
#pragma unroll
    for (int k = 0; k < F::kWords; k++) w[k] = L.w[k][lane];
    return F::from_words(w);
}

// Staple sum of U_MU(x) with the shared links from LDS.  XROWS: the tile is
// whole x-rows (n.xp, n.xm are the lanes of x+1, x-1), so the x-neighbours of
// the staged links are staged too.  Same operands and products, in the same
// order, as staple_sum (flow_body.hpp).
template <prec P, int MU, bool XROWS>
__device__ inline mat3<typename storage<P>::real> staple_sum_lds(
    const typename storage<P>::field& f, unsigned s, const nbr& n, const lds_link<P>* own,
    const lds_link<P>* dn, unsigned lane) {
    using T = typename storage<P>::real;
    mat3<T> st;
    const unsigned xmu = s + n.up[MU];
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == MU) continue;
        const unsigned xnu = s + n.up[nu];
        const unsigned xmnu = s + n.dn[nu];
        const unsigned xmunu = xmu + n.dn[nu];
        {  // U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag; the first staple sets the sum
            const mat3<T> c = (XROWS && nu == 0) ? from_lds<P>(own[MU], n.xp) : f.link(MU, xnu);
            const mat3<T> b = mul_su3(from_lds<P>(own[nu], lane), c);
            const mat3<T> a = (XROWS && MU == 0) ? from_lds<P>(own[nu], n.xp) : f.link(nu, xmu);
            sum_su3_rows(st, mul_nd_rows(a, b), nu == (MU == 0 ? 1 : 0));
        }
        {  // (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
            const mat3<T> d = (XROWS && nu == 0) ? from_lds<P>(own[MU], n.xm) : f.link(MU, xmnu);
            const mat3<T> e = (XROWS && MU == 0) ? from_lds<P>(dn[nu], n.xp) : f.link(nu, xmunu);
            const mat3<T> b = mul_su3(d, e);
            const mat3<T> v =
                (XROWS && nu == 0) ? from_lds<P>(own[0], n.xm) : from_lds<P>(dn[nu], lane);
            add_su3_rows(st, mul_dn_rows(b, v));
        }
        // keep the next plane's loads below this plane's arithmetic (FP64; the
        // FP32-mixed code is scheduled freely, as in the plain kernel)
        if constexpr (P == prec::f64) __builtin_amdgcn_sched_barrier(0);
    }
    return st;
}

template <prec P, int STAGE, int MU, bool XROWS>
__device__ inline void stage_lds_body(const typename storage<P>::field& f,
                                      typename storage<P>::vec* __restrict__ Uout,
                                      typename storage<P>::avec* __restrict__ acc, unsigned s,
                                      const nbr& n, const lds_link<P>* own, const lds_link<P>* dn,
                                      unsigned lane, double cz, double ca, bool& warn) {
    using T = typename storage<P>::real;
    const mat3<T> st = staple_sum_lds<P, MU, XROWS>(f, s, n, own, dn, lane);
    const mat3<T> u = from_lds<P>(own[MU], lane);
    const alg8<T> z = neg_proj_ta(mul(u, st));
    stage_update<P, STAGE, MU>(z, u, f.V, Uout, acc, s, cz, ca, warn);
}

// Wavefronts per SIMD: FP64 3 (its registers and 48 KiB of LDS per
// workgroup), FP32-mixed 6 (24 KiB of LDS per workgroup).
template <prec P>
constexpr int kLdsStageWaves = P == prec::f64 ? 3 : 6;

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
    // Stage U_mu(x) and U_nu(x - nu).  With whole x-rows dn[0] is own[0] at
    // x-1 and is not staged; wavefront 0, which loads the fewest links later,
    // stages dn[1..3].  Otherwise wavefront mu stages dn[mu].
    stage_to_lds<P>(Uin, g.V, mu, s, own[mu]);
    if constexpr (XROWS) {
        if (mu == 0)
#pragma unroll
            for (int nu = 1; nu < 4; nu++) stage_to_lds<P>(Uin, g.V, nu, s + n.dn[nu], dn[nu]);
    } else {
        // n.dn[mu] by selects: indexed by the run-time mu the array would
        // go to scratch memory (44 bytes per lane written and read back)
        const int dmu = mu == 0 ? n.dn[0] : mu == 1 ? n.dn[1] : mu == 2 ? n.dn[2] : n.dn[3];
        stage_to_lds<P>(Uin, g.V, mu, s + dmu, dn[mu]);
    }
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
