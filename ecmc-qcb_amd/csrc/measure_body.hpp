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

// Products whose right operand is streamed from memory a row at a time:
//   R  = A B        (rows k of B)        R += A B      (rows k of B)
//   R  = A B^dag    (rows j of B)        R += A B^dag  (rows j of B)
//   R  = A^dag B    (rows k of A and B)  R  = M B      (rows i of M, B whole)
template <class F>
__device__ inline void acc_mul_s(mat3<typename F::real>& R, const mat3<typename F::real>& A,
                                 const F& f, int mu, unsigned s) {
    using T = typename F::real;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        cpx<T> b[3];
        f.row(mu, s, k, b);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) R.e[i][j] = madd(R.e[i][j], A.e[i][k], b[j]);
    }
}
template <class F>
__device__ inline void acc_mul_nd_s(mat3<typename F::real>& R, const mat3<typename F::real>& A,
                                    const F& f, int mu, unsigned s) {
    using T = typename F::real;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        cpx<T> b[3];
        f.row(mu, s, j, b);
#pragma unroll
        for (int i = 0; i < 3; i++)
            R.e[i][j] = madd_c(madd_c(madd_c(R.e[i][j], A.e[i][0], b[0]), A.e[i][1], b[1]), A.e[i][2],
                               b[2]);
    }
}
template <class F>
__device__ inline mat3<typename F::real> mul_rows_s(const F& f, int mu, unsigned s,
                                                    const mat3<typename F::real>& B) {
    using T = typename F::real;
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        cpx<T> a[3];
        f.row(mu, s, i, a);
#pragma unroll
        for (int j = 0; j < 3; j++)
            R.e[i][j] = madd(madd(a[0] * B.e[0][j], a[1], B.e[1][j]), a[2], B.e[2][j]);
    }
    return R;
}
template <class F>
__device__ inline mat3<typename F::real> mul_dn_s(const F& f, int mua, unsigned sa, int mub,
                                                  unsigned sb) {
    using T = typename F::real;
    mat3<T> R = zero3<T>();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        cpx<T> a[3], b[3];
        f.row(mua, sa, k, a);
        f.row(mub, sb, k, b);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) R.e[i][j] = madd_cl(R.e[i][j], a[i], b[j]);
    }
    return R;
}

// The four clover leaves with streamed operands (FP64 storage): at most the
// running sum, two products and one whole link are live.
template <class F, int MU, int NU>
__device__ inline mat3<typename F::real> clover_leaves_streamed(
    const F& f, unsigned s, unsigned xpmu, unsigned xpnu, unsigned xmmu, unsigned xmnu,
    unsigned xmmupnu, unsigned xmmumnu, unsigned xpmumnu, double& plaq) {
    using T = typename F::real;
    mat3<T> C = zero3<T>();
    {  // U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
        const mat3<T> a = mul_rows_s(f, MU, s, f.link(NU, xpmu));
        mat3<T> b = zero3<T>();
        acc_mul_nd_s(b, a, f, MU, xpnu);
        acc_mul_nd_s(C, b, f, NU, s);
        plaq += (double)(C.e[0][0].r + C.e[1][1].r + C.e[2][2].r);
    }
    {  // U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
        mat3<T> a = zero3<T>();
        acc_mul_nd_s(a, f.link(NU, s), f, MU, xmmupnu);
        mat3<T> b = zero3<T>();
        acc_mul_nd_s(b, a, f, NU, xmmu);
        acc_mul_s(C, b, f, MU, xmmu);
    }
    {  // U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
        //  = (U_nu(x-mu-nu) U_mu(x-mu))^dag U_mu(x-mu-nu) U_nu(x-nu)
        const mat3<T> a = mul_rows_s(f, NU, xmmumnu, f.link(MU, xmmu));
        mat3<T> d = zero3<T>();
#pragma unroll
        for (int k = 0; k < 3; k++) {  // d = a^dag U_mu(x-mu-nu)
            cpx<T> b[3];
            f.row(MU, xmmumnu, k, b);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) d.e[i][j] = madd_cl(d.e[i][j], a.e[k][i], b[j]);
        }
        acc_mul_s(C, d, f, NU, xmnu);
    }
    {  // U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag
        const mat3<T> a = mul_dn_s(f, NU, xmnu, MU, xmnu);
        mat3<T> b = zero3<T>();
        acc_mul_s(b, a, f, NU, xpmumnu);
        acc_mul_nd_s(C, b, f, MU, s);
    }
    return C;
}

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
    if constexpr (F::stream_rows) {
        C = clover_leaves_streamed<F, MU, NU>(f, s, xpmu, xpnu, xmmu, xmnu, xmmupnu, xmmumnu,
                                              xpmumnu, plaq);
    } else if constexpr (QCB_SU3_PRODUCTS) {
        // every partial product of a leaf is in SU(3): two rows + cross product
        {
            const mat3<T> b = mul_nd_su3(mul_su3(f.link(MU, s), f.link(NU, xpmu)), f.link(MU, xpnu));
            C = mul_nd_su3(b, f.link(NU, s));
            plaq += (double)(C.e[0][0].r + C.e[1][1].r + C.e[2][2].r);
        }
        {
            const mat3<T> b =
                mul_nd_su3(mul_nd_su3(f.link(NU, s), f.link(MU, xmmupnu)), f.link(NU, xmmu));
            add_to(C, mul_su3(b, f.link(MU, xmmu)));
        }
        add_to(C, mul_dn_su3(mul_su3(f.link(NU, xmmumnu), f.link(MU, xmmu)),
                             mul_su3(f.link(MU, xmmumnu), f.link(NU, xmnu))));
        {
            const mat3<T> b = mul_su3(mul_dn_su3(f.link(NU, xmnu), f.link(MU, xmnu)),
                                      f.link(NU, xpmumnu));
            add_to(C, mul_nd_su3(b, f.link(MU, s)));
        }
    } else {
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
    if constexpr (QCB_SU3_PRODUCTS && !F::stream_rows) {
        const mat3<T> b = mul_nd_su3(mul_su3(f.link(MU, s), f.link(NU, xpmu)), f.link(MU, xpnu));
        plaq += (double)re_tr_nd(b, f.link(NU, s));
    } else {
        const mat3<T> a = mul(f.link(MU, s), f.link(NU, xpmu));
        const mat3<T> b = mul_nd(a, f.link(MU, xpnu));
        plaq += (double)re_tr_nd(b, f.link(NU, s));
    }
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

}  // namespace qcb
