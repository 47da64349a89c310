// stage_lds.hpp -- device code of the RK3 stage with the links that the four
// direction wavefronts of a tile share staged in LDS, shared by the
// single-lattice kernel (stage_lds.hip) and the ensemble kernel (batch.hip).
#pragma once
#include "flow_body.hpp"

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

// Stage U_mu(x) and U_nu(x - nu) of the tile for wavefront mu.  With whole
// x-rows dn[0] is own[0] at x-1 and is not staged; wavefront 0, which loads the
// fewest links later, stages dn[1..3].  Otherwise wavefront mu stages dn[mu],
// its offset n.dn[mu] picked by selects: indexed by the run-time mu, the array
// would go to scratch memory (44 bytes per lane written and read back).
template <prec P, bool XROWS>
__device__ inline void stage_tile_to_lds(const typename storage<P>::vec* __restrict__ U,
                                         unsigned V, int mu, unsigned s, const nbr& n,
                                         lds_link<P>* own, lds_link<P>* dn) {
    stage_to_lds<P>(U, V, mu, s, own[mu]);
    if constexpr (XROWS) {
        if (mu == 0)
#pragma unroll
            for (int nu = 1; nu < 4; nu++) stage_to_lds<P>(U, V, nu, s + n.dn[nu], dn[nu]);
    } else {
        const int dmu = mu == 0 ? n.dn[0] : mu == 1 ? n.dn[1] : mu == 2 ? n.dn[2] : n.dn[3];
        stage_to_lds<P>(U, V, mu, s + dmu, dn[mu]);
    }
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

}  // namespace qcb
