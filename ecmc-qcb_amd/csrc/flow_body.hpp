// flow_body.hpp -- per-link device code of the fused RK3 stage, shared by the
// single-lattice kernels (flow.hip) and the ensemble kernels (batch.hip).
#pragma once
#include <type_traits>

#include "fields.hpp"

namespace qcb {

// Scheduling barriers in the FP64 staple loop: 1 = one per plane (measured
// 1-2 % faster at 32^4 and 48^4), 2 = also between the upper and lower
// staple of a plane, 0 = none.  FP32 code is VALU-bound and has none.
#ifndef QCB_SCHED_SPLIT
#define QCB_SCHED_SPLIT 1
#endif

// Upper and lower staple of U_mu(x) in the plane (mu, nu), each in SU(3):
//   upper = U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag
//   lower = (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
// (the two terms of GaugeField::compute_staple, src/gauge/GaugeField.cpp:74-97)
// *_rows: rows 0-1 only (the staple is in SU(3); its third row is rebuilt
// where it is needed, su3.hpp).
template <class F, int MU>
__device__ inline mat3<typename F::real> upper_rows(const F& f, unsigned s, const nbr& n,
                                                    int nu) {
    const unsigned xmu = s + n.up[MU], xnu = s + n.up[nu];
    return mul_nd_rows(f.link(nu, xmu), mul_su3(f.link(nu, s), f.link(MU, xnu)));
}
template <class F, int MU>
__device__ inline mat3<typename F::real> lower_rows(const F& f, unsigned s, const nbr& n,
                                                    int nu) {
    const unsigned xmnu = s + n.dn[nu], xmunu = s + n.up[MU] + n.dn[nu];
    return mul_dn_rows(mul_su3(f.link(MU, xmnu), f.link(nu, xmunu)), f.link(nu, xmnu));
}
template <class F, int MU>
__device__ inline mat3<typename F::real> upper_staple(const F& f, unsigned s, const nbr& n,
                                                      int nu) {
    mat3<typename F::real> w = upper_rows<F, MU>(f, s, n, nu);
    complete_su3(w);
    return w;
}
template <class F, int MU>
__device__ inline mat3<typename F::real> lower_staple(const F& f, unsigned s, const nbr& n,
                                                      int nu) {
    mat3<typename F::real> w = lower_rows<F, MU>(f, s, n, nu);
    complete_su3(w);
    return w;
}

// The link words of lane src (the words a load at that lane's site returns).
template <class F>
__device__ inline void shfl_words(const typename F::word (&w)[F::kWords], int src,
                                  typename F::word (&out)[F::kWords]) {
    // a word is a double2 (FP64) or a float4 (FP32 two-row form)
    constexpr bool four = sizeof(typename F::word) == 4 * sizeof(typename F::real);
#pragma unroll
    for (int k = 0; k < F::kWords; k++) {
        out[k].x = __shfl(w[k].x, src, 64);
        out[k].y = __shfl(w[k].y, src, 64);
        if constexpr (four) {
            out[k].z = __shfl(w[k].z, src, 64);
            out[k].w = __shfl(w[k].w, src, 64);
        }
    }
}

// Receives the two staples of every plane as the staple sum forms them, the
// upper one first (the fused measurement builds three clover leaves from
// them); the default does nothing.
struct no_staple_hook {
    template <class M>
    __device__ void upper(int, const M&) {}
    template <class M>
    __device__ void lower(int, const M&) {}
};

// Adds the staple w (rows 0-1 given) to the staple sum st (st = w for the
// first), and hands the whole staple to the hook.  The sum forms the third
// row of w only inside its own multiply-add chain (add_su3_rows); a hook that
// needs it gets it rebuilt on its own, so every kernel sums the staples with
// the same roundings, hook or not.
template <bool UPPER, class Hook, typename T>
__device__ inline void take_staple(mat3<T>& st, const mat3<T>& w, bool first, int nu,
                                   Hook&& hook) {
    sum_su3_rows(st, w, first);
    if constexpr (!std::is_same<typename std::decay<Hook>::type, no_staple_hook>::value) {
        mat3<T> full = w;
        complete_su3(full);
        if constexpr (UPPER)
            hook.upper(nu, full);
        else
            hook.lower(nu, full);
    }
}

// Staple sum of U_mu(x):
//   sum_{nu != mu} U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag
//                + (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
// With SU(3) products (the default) each term is upper_staple / lower_staple.
// Without them, for FP64 one operand of every product is streamed from memory
// a row at a time, so only the accumulator, one product and one full operand
// are live; FP32 matrices are half the registers and are loaded whole.  The
// hook is called only on the SU(3)-product path.  XROWS: the wavefront holds
// whole x-rows (x_rows_in_wave, lattice.hpp; n.xp and n.xm set), so the links
// of the x-neighbours come from lane shuffles.
template <class F, int MU, bool XROWS = false, class Hook = no_staple_hook>
__device__ inline mat3<typename F::real> staple_sum(const F& f, unsigned s, const nbr& n,
                                                    Hook&& hook = Hook{}) {
    using T = typename F::real;
    // zero for the streamed-row and full-product paths; the SU(3)-product
    // path sets it from its first staple
    mat3<T> st = zero3<T>();
    const unsigned xmu = s + n.up[MU];
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == MU) continue;
        const unsigned xnu = s + n.up[nu];
        const unsigned xmnu = s + n.dn[nu];
        const unsigned xmunu = xmu + n.dn[nu];
        if constexpr (F::stream_rows) {
            {  // T = U_nu(x) U_mu(x+nu); st += U_nu(x+mu) T^dag
                const mat3<T> B = f.link(MU, xnu);
                mat3<T> P;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    cpx<T> a[3];
                    f.row(nu, s, i, a);
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        P.e[i][j] = madd(madd(a[0] * B.e[0][j], a[1], B.e[1][j]), a[2], B.e[2][j]);
                }
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    cpx<T> c[3];
                    f.row(nu, xmu, i, c);
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        st.e[i][j] = madd_c(madd_c(madd_c(st.e[i][j], c[0], P.e[j][0]), c[1],
                                                   P.e[j][1]),
                                            c[2], P.e[j][2]);
                }
            }
            {  // T = U_mu(x-nu) U_nu(x+mu-nu); st += T^dag U_nu(x-nu)
                const mat3<T> B = f.link(nu, xmunu);
                mat3<T> P;
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    cpx<T> a[3];
                    f.row(MU, xmnu, i, a);
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        P.e[i][j] = madd(madd(a[0] * B.e[0][j], a[1], B.e[1][j]), a[2], B.e[2][j]);
                }
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    cpx<T> v[3];
                    f.row(nu, xmnu, k, v);  // row k of U_nu(x-nu)
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) st.e[i][j] = madd_cl(st.e[i][j], P.e[k][i], v[j]);
                }
            }
        } else if constexpr (QCB_SU3_PRODUCTS) {
            // Every product of links is in SU(3): form two rows, rebuild the
            // third.  The first staple (the upper one of the first plane) sets
            // the sum.
            const bool first = nu == (MU == 0 ? 1 : 0);
            if (XROWS && (MU == 0 || nu == 0)) {
                // The x-neighbours of this plane's links are in this wavefront:
                // shuffle them instead of loading them (same words, same
                // products as upper_rows / lower_rows).
                using W = typename F::word;
                if (MU == 0) {  // U_nu(x+mu) is lane xp's U_nu(x)
                    W a[F::kWords], ap[F::kWords];
                    f.load_words(nu, s, a);
                    shfl_words<F>(a, n.xp, ap);
                    take_staple<true>(st,
                                      mul_nd_rows(F::from_words(ap),
                                                  mul_su3(F::from_words(a), f.link(MU, xnu))),
                                      first, nu, hook);
                    take_staple<false>(st, lower_rows<F, MU>(f, s, n, nu), false, nu, hook);
                } else {  // nu == 0: U_mu(x+nu), U_mu(x-nu), U_nu(x-nu) from lanes
                    W own[F::kWords], a[F::kWords], t[F::kWords];
                    f.load_words(MU, s, own);
                    f.load_words(nu, s, a);
                    shfl_words<F>(own, n.xp, t);
                    take_staple<true>(st,
                                      mul_nd_rows(f.link(nu, xmu),
                                                  mul_su3(F::from_words(a), F::from_words(t))),
                                      first, nu, hook);
                    shfl_words<F>(own, n.xm, t);
                    const mat3<T> b = mul_su3(F::from_words(t), f.link(nu, xmunu));
                    shfl_words<F>(a, n.xm, t);
                    take_staple<false>(st, mul_dn_rows(b, F::from_words(t)), false, nu, hook);
                }
            } else {
                take_staple<true>(st, upper_rows<F, MU>(f, s, n, nu), first, nu, hook);
                // FP64 without x-row shuffles: the lower staple's loads stay
                // below the upper one's arithmetic too, which keeps the plain
                // and ensemble kernels within three wavefronts' registers
                // (150-158 VGPRs, against 168 and 12-23 spilled without)
                if constexpr (sizeof(T) == 8 && (QCB_SCHED_SPLIT >= 2 || !XROWS))
                    __builtin_amdgcn_sched_barrier(0);
                take_staple<false>(st, lower_rows<F, MU>(f, s, n, nu), false, nu, hook);
            }
        } else {
            {
                const mat3<T> a = mul(f.link(nu, s), f.link(MU, xnu));
                acc_mul_nd(st, f.link(nu, xmu), a);
            }
#if QCB_SCHED_SPLIT >= 2
            if constexpr (sizeof(T) == 8) __builtin_amdgcn_sched_barrier(0);
#endif
            {
                const mat3<T> b = mul(f.link(MU, xmnu), f.link(nu, xmunu));
                acc_mul_dn(st, b, f.link(nu, xmnu));
            }
        }
#if QCB_SCHED_SPLIT >= 1
        // keep the scheduler from hoisting the next plane's loads above this
        // one's arithmetic: fewer live registers, more resident wavefronts
        if constexpr (sizeof(T) == 8) __builtin_amdgcn_sched_barrier(0);
#endif
    }
    return st;
}

template <class F, int MU, bool XROWS = false, class Hook = no_staple_hook>
__device__ inline alg8<typename F::real> force_at(const F& f, unsigned s, const nbr& n,
                                                  mat3<typename F::real>& u,
                                                  Hook&& hook = Hook{}) {
    const mat3<typename F::real> st = staple_sum<F, MU, XROWS>(f, s, n, hook);
    u = f.link(MU, s);
    return neg_proj_ta(mul(u, st));
}

// The RK3 combination of the force z of U_mu(s) = u with the accumulator,
// the exponential and the update exp(X) u into Uout (compute_w1/w2/w3 and
// replace_force_0, gradient_flow.cpp:40-121).
template <prec P, int STAGE, int MU>
__device__ inline void stage_update(const alg8<typename storage<P>::real>& z,
                                    const mat3<typename storage<P>::real>& u, unsigned V,
                                    typename storage<P>::vec* __restrict__ Uout,
                                    typename storage<P>::avec* __restrict__ acc, unsigned s,
                                    double cz, double ca, bool& warn) {
    using S = storage<P>;
    using T = typename S::real;
    const T czr = (T)cz, car = (T)ca;
    alg8<T> x;
    if constexpr (STAGE == 1) {
#pragma unroll
        for (int k = 0; k < 8; k++) x.a[k] = czr * z.a[k];
        S::store_alg(acc, V, MU, s, z);
    } else {
        alg8<T> a = S::load_alg(acc, V, MU, s);
#pragma unroll
        for (int k = 0; k < 8; k++) x.a[k] = czr * z.a[k] - car * a.a[k];
        if constexpr (STAGE == 2) {
#pragma unroll
            for (int k = 0; k < 8; k++)
                a.a[k] = (T)(8.0 / 9.0) * z.a[k] - (T)(17.0 / 36.0) * a.a[k];
            S::store_alg(acc, V, MU, s, a);
        }
    }
    const mat3<T> e = exp_su3<T, double>(x, &warn);
    S::store_stage(Uout, V, MU, s, mul(e, u), STAGE);
}

template <prec P, int STAGE, int MU, bool XROWS = false, class Hook = no_staple_hook>
__device__ inline void stage_body(const typename storage<P>::field& f,
                                  typename storage<P>::vec* __restrict__ Uout,
                                  typename storage<P>::avec* __restrict__ acc, unsigned s,
                                  const nbr& n, double cz, double ca, bool& warn,
                                  Hook&& hook = Hook{}) {
    using T = typename storage<P>::real;
    mat3<T> u;
    const alg8<T> z = force_at<typename storage<P>::field, MU, XROWS>(f, s, n, u, hook);
    stage_update<P, STAGE, MU>(z, u, f.V, Uout, acc, s, cz, ca, warn);
}

// The site index through an empty asm statement: the compiler must assume it
// changed, so loads addressed from it are never merged with earlier loads.
__device__ __forceinline__ unsigned opaque_site(unsigned s) {
    asm volatile("" : "+v"(s));
    return s;
}

// One of several direction bodies run back to back by one wavefront.  Merged
// with the previous body's loads, the links the directions share would stay
// live across bodies and spill; through opaque_site and behind a scheduling
// barrier each body needs only the registers of a one-direction wavefront and
// reads the shared links again (from the L1).
template <prec P, int STAGE, int MU, bool XROWS>
__device__ __forceinline__ void stage_body_next(const typename storage<P>::field& f,
                                                typename storage<P>::vec* __restrict__ Uout,
                                                typename storage<P>::avec* __restrict__ acc,
                                                unsigned s, const nbr& n, double cz, double ca,
                                                bool& warn) {
    __builtin_amdgcn_sched_barrier(0);
    stage_body<P, STAGE, MU, XROWS>(f, Uout, acc, opaque_site(s), n, cz, ca, warn);
}

// The directions of site s that wavefront y (threadIdx.y) updates: with
// q = y mod 4/DPW, directions q*DPW .. q*DPW + DPW - 1 (threads per site =
// 4/DPW).
template <prec P, int STAGE, bool XROWS, int DPW>
__device__ __forceinline__ void stage_directions(unsigned y,
                                                 const typename storage<P>::vec* __restrict__ Uin,
                                                 typename storage<P>::vec* __restrict__ Uout,
                                                 typename storage<P>::avec* __restrict__ acc,
                                                 unsigned V, unsigned s, const nbr& n, double cz,
                                                 double ca, bool& w) {
    static_assert(DPW == 1 || DPW == 2 || DPW == 4, "1, 2 or 4 directions per wavefront");
    const auto f = storage<P>::view(Uin, V);
    if constexpr (DPW == 1) {
        switch (y) {
            case 0: stage_body<P, STAGE, 0, XROWS>(f, Uout, acc, s, n, cz, ca, w); break;
            case 1: stage_body<P, STAGE, 1, XROWS>(f, Uout, acc, s, n, cz, ca, w); break;
            case 2: stage_body<P, STAGE, 2, XROWS>(f, Uout, acc, s, n, cz, ca, w); break;
            default: stage_body<P, STAGE, 3, XROWS>(f, Uout, acc, s, n, cz, ca, w); break;
        }
    } else if constexpr (DPW == 2) {
        if (__builtin_amdgcn_readfirstlane((int)y) % 2 == 0) {
            stage_body_next<P, STAGE, 0, XROWS>(f, Uout, acc, s, n, cz, ca, w);
            stage_body_next<P, STAGE, 1, XROWS>(f, Uout, acc, s, n, cz, ca, w);
        } else {
            stage_body_next<P, STAGE, 2, XROWS>(f, Uout, acc, s, n, cz, ca, w);
            stage_body_next<P, STAGE, 3, XROWS>(f, Uout, acc, s, n, cz, ca, w);
        }
    } else {
        stage_body_next<P, STAGE, 0, XROWS>(f, Uout, acc, s, n, cz, ca, w);
        stage_body_next<P, STAGE, 1, XROWS>(f, Uout, acc, s, n, cz, ca, w);
        stage_body_next<P, STAGE, 2, XROWS>(f, Uout, acc, s, n, cz, ca, w);
        stage_body_next<P, STAGE, 3, XROWS>(f, Uout, acc, s, n, cz, ca, w);
    }
}

}  // namespace qcb
