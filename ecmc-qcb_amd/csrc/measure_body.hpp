// measure_body.hpp -- per-site device code of the clover / plaquette / Q / E
// measurement, shared by measure.hip and batch.hip.
#pragma once
#include "fields.hpp"

namespace qcb {

// Clover field strength F = (C - C^dag)/8 (trace kept, as the reference does):
// anti-hermitian, stored as the imaginary diagonal and the upper triangle.
template <typename T>
struct fclov {
    T d[3];
    cpx<T> o[3];  // (0,1), (0,2), (1,2)
};

// Sum of the four leaves of plane (MU,NU) at x, all anticlockwise, in the
// order and association of clover_site (observables_mpi.cpp:78-86).  Adds
// Re tr(leaf 0) to plaq.
template <class F, int MU, int NU>
__device__ inline fclov<typename F::real> clover(const F& f, unsigned s, const nbr& n,
                                                 double& plaq) {
    using T = typename F::real;
    const unsigned xpmu = s + n.up[MU], xpnu = s + n.up[NU];
    const unsigned xmmu = s + n.dn[MU], xmnu = s + n.dn[NU];
    const unsigned xmmupnu = xmmu + n.up[NU], xmmumnu = xmmu + n.dn[NU];
    const unsigned xpmumnu = xpmu + n.dn[NU];
    mat3<T> C;
    {  // U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
        const mat3<T> a = mul(f.link(MU, s), f.link(NU, xpmu));
        const mat3<T> b = mul_nd(a, f.link(MU, xpnu));
        C = mul_nd(b, f.link(NU, s));
        plaq += (double)(C.e[0][0].r + C.e[1][1].r + C.e[2][2].r);
    }
    {  // U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
        const mat3<T> a = mul_nd(f.link(NU, s), f.link(MU, xmmupnu));
        const mat3<T> b = mul_nd(a, f.link(NU, xmmu));
        acc_mul(C, b, f.link(MU, xmmu));
    }
    {  // U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
        const mat3<T> a = mul(f.link(NU, xmmumnu), f.link(MU, xmmu));
        const mat3<T> b = mul(f.link(MU, xmmumnu), f.link(NU, xmnu));
        acc_mul_dn(C, a, b);
    }
    {  // U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag
        const mat3<T> a = mul_dn(f.link(NU, xmnu), f.link(MU, xmnu));
        const mat3<T> b = mul(a, f.link(NU, xpmumnu));
        acc_mul_nd(C, b, f.link(MU, s));
    }
    fclov<T> Fc;
    const T e8 = (T)0.125;
#pragma unroll
    for (int k = 0; k < 3; k++) Fc.d[k] = (C.e[k][k].i + C.e[k][k].i) * e8;
    Fc.o[0] = e8 * (C.e[0][1] - conj(C.e[1][0]));
    Fc.o[1] = e8 * (C.e[0][2] - conj(C.e[2][0]));
    Fc.o[2] = e8 * (C.e[1][2] - conj(C.e[2][1]));
    return Fc;
}

template <class F, int MU, int NU>
__device__ inline void plaq_only(const F& f, unsigned s, const nbr& n, double& plaq) {
    using T = typename F::real;
    const unsigned xpmu = s + n.up[MU], xpnu = s + n.up[NU];
    const mat3<T> a = mul(f.link(MU, s), f.link(NU, xpmu));
    const mat3<T> b = mul_nd(a, f.link(MU, xpnu));
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

template <class F, int M1, int N1, int M2, int N2, int SIGN, bool CLOVER>
__device__ inline void pair_body(const F& f, unsigned s, const nbr& n, double& vp, double& ve,
                                 double& vq) {
    if constexpr (CLOVER) {
        const auto A = clover<F, M1, N1>(f, s, n, vp);
        const auto B = clover<F, M2, N2>(f, s, n, vp);
        ve += 0.5 * ((double)fnorm2(A) + (double)fnorm2(B));
        vq += SIGN * (double)re_tr_ab(A, B);
    } else {
        plaq_only<F, M1, N1>(f, s, n, vp);
        plaq_only<F, M2, N2>(f, s, n, vp);
    }
}

__device__ inline void neumaier_add(double& sum, double& comp, double x) {
    const double t = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
}

}  // namespace qcb
