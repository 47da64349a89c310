// measure_body.hpp -- per-site device code of the clover / plaquette / Q / E
// measurement, shared by measure.hip and batch.hip.
#pragma once
#include "flow_body.hpp"

namespace qcb {

// Clover field strength F = (C - C^dag)/8 (trace kept, as the reference does):
// anti-hermitian, stored as the imaginary diagonal and the upper triangle.
template <typename T>
struct fclov {
    T d[3];
    cpx<T> o[3];  // (0,1), (0,2), (1,2)
};

// Re tr of the plaquette U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
// (mean_plaquette_local, observables_mpi.cpp:11-35), for the plaquette-only
// measurement.
template <class F, int MU, int NU>
__device__ inline void plaq_only(const F& f, unsigned s, const nbr& n, double& plaq) {
    using T = typename F::real;
    const unsigned xpmu = s + n.up[MU], xpnu = s + n.up[NU];
    const mat3<T> b = mul_nd_su3(mul_su3(f.link(MU, s), f.link(NU, xpmu)), f.link(MU, xpnu));
    plaq += (double)re_tr_nd(b, f.link(NU, s));
}

// |F|_F^2 for anti-hermitian F
template <typename T>
__device__ inline T fnorm2(const fclov<T>& F) {
    return F.d[0] * F.d[0] + F.d[1] * F.d[1] + F.d[2] * F.d[2] +
           T(2) * (norm2(F.o[0]) + norm2(F.o[1]) + norm2(F.o[2]));
}
// Re tr(A B) for anti-hermitian A, B
template <typename T>
__device__ inline T re_tr_ab(const fclov<T>& A, const fclov<T>& B) {
    T s = A.d[0] * B.d[0] + A.d[1] * B.d[1] + A.d[2] * B.d[2];
    T o = T(0);
#pragma unroll
    for (int k = 0; k < 3; k++) o += A.o[k].r * B.o[k].r + A.o[k].i * B.o[k].i;
    return -(s + T(2) * o);
}

// ---- clover leaves from staples ----------------------------------------------
// Of the four leaves of plane (a, b) at x (clover_site, observables_mpi.cpp:
// 66-90), three are one SU(3) product away from the staples that the flow
// forms for the links at x:
//   leaf 0 = U_a(x) upper(a,b)            = (U_b(x) upper(b,a))^dag
//   leaf 1 = U_b(x) lower(b,a)
//   leaf 3 = (U_a(x) lower(a,b))^dag
//   leaf 2 = (U_b(x-a-b) U_a(x-a))^dag U_a(x-a-b) U_b(x-b)   (formed directly)
// where upper(mu,nu) / lower(mu,nu) are the staples of U_mu(x) in the plane
// (flow_body.hpp).  F = (C - C^dag)/8 is linear in the leaves, so each
// wavefront (direction mu) adds up the anti-hermitian parts of the leaves it
// forms for each of its three planes and leaves them in LDS; the two
// wavefronts of a plane then meet there.  The same code runs inside the fused
// RK3 stage 1 (where the staples are already formed) and in the standalone
// measurement, so both give bit-identical results.

// Plane (a, b), a < b: (0,1) 0, (0,2) 1, (0,3) 2, (1,2) 3, (1,3) 4, (2,3) 5.
__host__ __device__ constexpr int plane_of(int a, int b) { return a == 0 ? b - 1 : a + b; }
// Wavefront that forms leaf 0 and leaf 2 of each plane.  Every wavefront forms
// one leaf (1 or 3) per plane from its lower staples; with these owners the
// four wavefronts form 8, 8, 10 and 10 SU(3) products per site.
__host__ __device__ constexpr int leaf0_wave(int p) {
    return p == 2 ? 3 : (p == 3 || p == 4) ? 1 : p == 5 ? 2 : 0;
}
__host__ __device__ constexpr int leaf2_wave(int p) {
    return p == 0 ? 1 : (p == 1 || p == 3) ? 2 : p == 2 ? 0 : 3;
}
// LDS slot of wavefront mu's part of the plane (mu, nu): three per wavefront.
__host__ __device__ constexpr int clover_slot(int mu, int nu) { return 3 * mu + (nu < mu ? nu : nu - 1); }
constexpr int kCloverSlots = 12;

// F += sign (M - M^dag)/8
template <typename T>
__device__ inline void add_ah8(fclov<T>& F, const mat3<T>& M, T sign) {
    const T e = sign * (T)0.125;
#pragma unroll
    for (int k = 0; k < 3; k++) F.d[k] += e * (M.e[k][k].i + M.e[k][k].i);
    F.o[0] = F.o[0] + e * (M.e[0][1] - conj(M.e[1][0]));
    F.o[1] = F.o[1] + e * (M.e[0][2] - conj(M.e[2][0]));
    F.o[2] = F.o[2] + e * (M.e[1][2] - conj(M.e[2][1]));
}

template <class F>
__device__ inline mat3<typename F::real> clover_leaf2(const F& f, unsigned s, const nbr& n, int a,
                                                      int b) {
    const unsigned xma = s + n.dn[a], xmb = s + n.dn[b], y = xma + n.dn[b];
    return mul_dn_su3(mul_su3(f.link(b, y), f.link(a, xma)), mul_su3(f.link(a, y), f.link(b, xmb)));
}

// Staple hook of wavefront MU.  For each plane (MU, nu) it forms the leaves
// 0 (if it owns it) and 1 or 3 from the staples as they arrive, adds their
// anti-hermitian parts and writes them to slot clover_slot(MU, nu) of lds
// (layout [slot][9 reals][64 lanes]); finish() then adds the leaves 2 it owns,
// formed from their own links, once the staples' registers are free.  The
// real traces of the leaf-0 products go to plaq.
template <class F, int MU>
struct clover_parts {
    using T = typename F::real;
    const F& f;
    unsigned s;
    const nbr& n;
    mat3<T> u;  // U_mu(x); rows 0-1 are read
    T* lds;
    unsigned lane;
    double plaq;
    fclov<T> c;  // the current plane's part

    __device__ static constexpr int plane(int nu) {
        return plane_of(MU < nu ? MU : nu, MU < nu ? nu : MU);
    }
    __device__ T* slot(int nu) const { return lds + (size_t)clover_slot(MU, nu) * 9 * 64 + lane; }

    __device__ void upper(int nu, const mat3<T>& wu) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            c.d[k] = T(0);
            c.o[k] = {T(0), T(0)};
        }
        if (leaf0_wave(plane(nu)) == MU) {
            const mat3<T> m = mul_su3(u, wu);  // leaf 0 (MU < nu) or its adjoint
            add_ah8(c, m, MU < nu ? T(1) : T(-1));
            plaq += (double)(m.e[0][0].r + m.e[1][1].r + m.e[2][2].r);
        }
    }
    __device__ void lower(int nu, const mat3<T>& wd) {
        // leaf 3 = (u lower)^dag when MU < nu, leaf 1 = u lower when MU > nu
        add_ah8(c, mul_su3(u, wd), MU < nu ? T(-1) : T(1));
        T* q = slot(nu);
#pragma unroll
        for (int k = 0; k < 3; k++) q[k * 64] = c.d[k];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            q[(3 + 2 * k) * 64] = c.o[k].r;
            q[(4 + 2 * k) * 64] = c.o[k].i;
        }
    }
    __device__ void finish() {
#pragma unroll
        for (int nu = 0; nu < 4; nu++) {
            if (nu == MU || leaf2_wave(plane(nu)) != MU) continue;
            fclov<T> l;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                l.d[k] = T(0);
                l.o[k] = {T(0), T(0)};
            }
            add_ah8(l, clover_leaf2(f, s, n, MU < nu ? MU : nu, MU < nu ? nu : MU), T(1));
            T* q = slot(nu);
#pragma unroll
            for (int k = 0; k < 3; k++) q[k * 64] += l.d[k];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                q[(3 + 2 * k) * 64] += l.o[k].r;
                q[(4 + 2 * k) * 64] += l.o[k].i;
            }
        }
    }
};

// Standalone measurement work of wavefront MU: the staples its leaves need.
template <class F, int MU>
__device__ inline double clover_wave(const F& f, unsigned s, const nbr& n,
                                     typename F::real* lds, unsigned lane) {
    clover_parts<F, MU> h{f, s, n, f.link(MU, s), lds, lane, 0.0, {}};
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == MU) continue;
        if (leaf0_wave(h.plane(nu)) == MU)
            h.upper(nu, upper_staple<F, MU>(f, s, n, nu));
        else
            h.upper(nu, mat3<typename F::real>{});  // clears the plane's part only
        h.lower(nu, lower_staple<F, MU>(f, s, n, nu));
    }
    h.finish();
    return h.plaq;
}

// F of plane (a, b) from the two wavefronts' parts, wavefront a's first.
template <typename T>
__device__ inline fclov<T> clover_from_lds(const T* lds, unsigned lane, int a, int b) {
    const T* pa = lds + (size_t)clover_slot(a, b) * 9 * 64 + lane;
    const T* pb = lds + (size_t)clover_slot(b, a) * 9 * 64 + lane;
    fclov<T> F;
#pragma unroll
    for (int k = 0; k < 3; k++) F.d[k] = pa[k * 64] + pb[k * 64];
#pragma unroll
    for (int k = 0; k < 3; k++)
        F.o[k] = {pa[(3 + 2 * k) * 64] + pb[(3 + 2 * k) * 64],
                  pa[(4 + 2 * k) * 64] + pb[(4 + 2 * k) * 64]};
    return F;
}

// After the workgroup barrier: wavefront w < 3 takes the plane pair
// {(0,1),(2,3)}, {(0,2),(1,3)} or {(0,3),(1,2)} and adds its E and Q terms
// (local_q_e_clover, observables_mpi.cpp:93-112).
template <typename T>
__device__ inline void clover_pair_terms(const T* lds, unsigned lane, int w, double& ve,
                                         double& vq) {
    if (w >= 3) return;
    const int b = w + 1;                  // partner of 0
    const int c = w == 0 ? 2 : 1, d = w == 2 ? 2 : 3;
    const fclov<T> A = clover_from_lds(lds, lane, 0, b);
    const fclov<T> B = clover_from_lds(lds, lane, c, d);
    ve += 0.5 * ((double)fnorm2(A) + (double)fnorm2(B));
    vq += (w == 1 ? -1.0 : 1.0) * (double)re_tr_ab(A, B);
}

}  // namespace qcb
