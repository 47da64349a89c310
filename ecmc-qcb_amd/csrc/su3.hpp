// su3.hpp -- device-inline 3x3 complex algebra for SU(3) links and su(3) elements.
//
// Replaces the Eigen fixed-size products the reference uses for every 3x3
// operation on the flow path (src/su3/utils.h:12, src/gauge/GaugeField.h:16).
// Everything is templated on the real type so the FP64 path and the FP32-mixed
// path share one implementation; all of it runs on the vector ALUs (3x3
// complex products are not a dense contraction worth an MFMA tile).
#pragma once
#include <hip/hip_runtime.h>

namespace qcb {

template <typename T>
struct cpx {
    T r, i;
};

template <typename T>
__host__ __device__ inline cpx<T> mk(T r, T i) {
    return cpx<T>{r, i};
}
template <typename T>
__host__ __device__ inline cpx<T> operator+(cpx<T> a, cpx<T> b) {
    return {a.r + b.r, a.i + b.i};
}
template <typename T>
__host__ __device__ inline cpx<T> operator-(cpx<T> a, cpx<T> b) {
    return {a.r - b.r, a.i - b.i};
}
template <typename T>
__host__ __device__ inline cpx<T> operator*(cpx<T> a, cpx<T> b) {
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename T>
__host__ __device__ inline cpx<T> operator*(T s, cpx<T> a) {
    return {s * a.r, s * a.i};
}
template <typename T>
__host__ __device__ inline cpx<T> conj(cpx<T> a) {
    return {a.r, -a.i};
}
// acc + a*b
template <typename T>
__host__ __device__ inline cpx<T> madd(cpx<T> acc, cpx<T> a, cpx<T> b) {
    return {acc.r + a.r * b.r - a.i * b.i, acc.i + a.r * b.i + a.i * b.r};
}
// acc + a*conj(b)
template <typename T>
__host__ __device__ inline cpx<T> madd_c(cpx<T> acc, cpx<T> a, cpx<T> b) {
    return {acc.r + a.r * b.r + a.i * b.i, acc.i + a.i * b.r - a.r * b.i};
}
// acc + conj(a)*b
template <typename T>
__host__ __device__ inline cpx<T> madd_cl(cpx<T> acc, cpx<T> a, cpx<T> b) {
    return {acc.r + a.r * b.r + a.i * b.i, acc.i + a.r * b.i - a.i * b.r};
}
template <typename T>
__host__ __device__ inline T norm2(cpx<T> a) {
    return a.r * a.r + a.i * a.i;
}

// Row-major 3x3 complex matrix.
template <typename T>
struct mat3 {
    cpx<T> e[3][3];
};

// Traceless anti-hermitian 3x3 matrix (an element of su(3)) in 8 reals:
//   [[ i d0,   p,      q    ],
//    [ -p*,    i d1,   r    ],
//    [ -q*,    -r*,    i d2 ]],  d2 = -d0 - d1.
// a[0]=d0 a[1]=d1 a[2..3]=p a[4..5]=q a[6..7]=r.
template <typename T>
struct alg8 {
    T a[8];
};

template <typename T>
__host__ __device__ inline mat3<T> zero3() {
    mat3<T> m;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) m.e[i][j] = {T(0), T(0)};
    return m;
}

// R = A B
template <typename T>
__host__ __device__ inline mat3<T> mul(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = A.e[i][0] * B.e[0][j];
            s = madd(s, A.e[i][1], B.e[1][j]);
            s = madd(s, A.e[i][2], B.e[2][j]);
            R.e[i][j] = s;
        }
    return R;
}
// R = A B^dag
template <typename T>
__host__ __device__ inline mat3<T> mul_nd(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = A.e[i][0] * conj(B.e[j][0]);
            s = madd_c(s, A.e[i][1], B.e[j][1]);
            s = madd_c(s, A.e[i][2], B.e[j][2]);
            R.e[i][j] = s;
        }
    return R;
}
// R = A^dag B
template <typename T>
__host__ __device__ inline mat3<T> mul_dn(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = conj(A.e[0][i]) * B.e[0][j];
            s = madd_cl(s, A.e[1][i], B.e[1][j]);
            s = madd_cl(s, A.e[2][i], B.e[2][j]);
            R.e[i][j] = s;
        }
    return R;
}
// R += A B
template <typename T>
__host__ __device__ inline void acc_mul(mat3<T>& R, const mat3<T>& A, const mat3<T>& B) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = R.e[i][j];
            s = madd(s, A.e[i][0], B.e[0][j]);
            s = madd(s, A.e[i][1], B.e[1][j]);
            s = madd(s, A.e[i][2], B.e[2][j]);
            R.e[i][j] = s;
        }
}
// R += A B^dag
template <typename T>
__host__ __device__ inline void acc_mul_nd(mat3<T>& R, const mat3<T>& A, const mat3<T>& B) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = R.e[i][j];
            s = madd_c(s, A.e[i][0], B.e[j][0]);
            s = madd_c(s, A.e[i][1], B.e[j][1]);
            s = madd_c(s, A.e[i][2], B.e[j][2]);
            R.e[i][j] = s;
        }
}
// R += A^dag B
template <typename T>
__host__ __device__ inline void acc_mul_dn(mat3<T>& R, const mat3<T>& A, const mat3<T>& B) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = R.e[i][j];
            s = madd_cl(s, A.e[0][i], B.e[0][j]);
            s = madd_cl(s, A.e[1][i], B.e[1][j]);
            s = madd_cl(s, A.e[2][i], B.e[2][j]);
            R.e[i][j] = s;
        }
}
// ---- products of two SU(3) matrices --------------------------------------------
// The product of two SU(3) matrices is in SU(3), so its third row is
// conj(row0 x row1) (U^dag = adj U when det U = 1).  The *_su3 products form
// rows 0 and 1 (18 complex multiply-adds) and the cross product (6 complex
// products) instead of the 27 multiply-adds of the third row, and they read
// only rows 0-1 of a left operand that enters by rows, which lets the caller
// skip rebuilding that operand's third row.  Equal to the full product up to
// rounding of the order of the operands' distance from SU(3).
template <typename T>
__host__ __device__ inline void complete_su3(mat3<T>& m) {
    m.e[2][0] = conj(m.e[0][1] * m.e[1][2] - m.e[0][2] * m.e[1][1]);
    m.e[2][1] = conj(m.e[0][2] * m.e[1][0] - m.e[0][0] * m.e[1][2]);
    m.e[2][2] = conj(m.e[0][0] * m.e[1][1] - m.e[0][1] * m.e[1][0]);
}
// A B, A and B in SU(3) (reads rows 0-1 of A)
template <typename T>
__host__ __device__ inline mat3<T> mul_su3(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = A.e[i][0] * B.e[0][j];
            s = madd(s, A.e[i][1], B.e[1][j]);
            s = madd(s, A.e[i][2], B.e[2][j]);
            R.e[i][j] = s;
        }
    complete_su3(R);
    return R;
}
// Rows 0-1 of A B^dag (row 2 of the result is left unset; reads rows 0-1 of A)
template <typename T>
__host__ __device__ inline mat3<T> mul_nd_rows(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = A.e[i][0] * conj(B.e[j][0]);
            s = madd_c(s, A.e[i][1], B.e[j][1]);
            s = madd_c(s, A.e[i][2], B.e[j][2]);
            R.e[i][j] = s;
        }
    return R;
}
// Rows 0-1 of A^dag B (row 2 of the result is left unset)
template <typename T>
__host__ __device__ inline mat3<T> mul_dn_rows(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cpx<T> s = conj(A.e[0][i]) * B.e[0][j];
            s = madd_cl(s, A.e[1][i], B.e[1][j]);
            s = madd_cl(s, A.e[2][i], B.e[2][j]);
            R.e[i][j] = s;
        }
    return R;
}
// A B^dag, A and B in SU(3) (reads rows 0-1 of A)
template <typename T>
__host__ __device__ inline mat3<T> mul_nd_su3(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R = mul_nd_rows(A, B);
    complete_su3(R);
    return R;
}
// A^dag B, A and B in SU(3)
template <typename T>
__host__ __device__ inline mat3<T> mul_dn_su3(const mat3<T>& A, const mat3<T>& B) {
    mat3<T> R = mul_dn_rows(A, B);
    complete_su3(R);
    return R;
}
// Sums of SU(3) matrices given by their rows 0-1 (W's row 2 is conj(row0 x row1)):
//   set_su3_rows: S = W;   add_su3_rows: S += W,
// the cross product entering S's third row inside its multiply-add chain, so
// the third row of W is never formed and added on its own.
template <typename T>
__host__ __device__ inline void set_su3_rows(mat3<T>& S, const mat3<T>& W) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S.e[i][j] = W.e[i][j];
    complete_su3(S);
}
// acc + conj(a b - c d), with acc first in the chain
template <typename T>
__host__ __device__ inline cpx<T> add_conj_cross(cpx<T> acc, cpx<T> a, cpx<T> b, cpx<T> c,
                                                 cpx<T> d) {
    return {acc.r + a.r * b.r - a.i * b.i - c.r * d.r + c.i * d.i,
            acc.i - a.r * b.i - a.i * b.r + c.r * d.i + c.i * d.r};
}
template <typename T>
__host__ __device__ inline void add_su3_rows(mat3<T>& S, const mat3<T>& W) {
#if defined(QCB_TRY_CLASSIC)
    mat3<T> F = W;
    complete_su3(F);
    add_to(S, F);
    return;
#endif
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S.e[i][j] = S.e[i][j] + W.e[i][j];
    S.e[2][0] = add_conj_cross(S.e[2][0], W.e[0][1], W.e[1][2], W.e[0][2], W.e[1][1]);
    S.e[2][1] = add_conj_cross(S.e[2][1], W.e[0][2], W.e[1][0], W.e[0][0], W.e[1][2]);
    S.e[2][2] = add_conj_cross(S.e[2][2], W.e[0][0], W.e[1][1], W.e[0][1], W.e[1][0]);
}
template <typename T>
__host__ __device__ inline void sum_su3_rows(mat3<T>& S, const mat3<T>& W, bool first) {
    if (first)
        set_su3_rows(S, W);
    else
        add_su3_rows(S, W);
}
template <typename T>
__host__ __device__ inline void add_to(mat3<T>& R, const mat3<T>& A) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) R.e[i][j] = R.e[i][j] + A.e[i][j];
}

// Re tr(A B^dag) = sum_ij Re(A_ij conj(B_ij))
template <typename T>
__host__ __device__ inline T re_tr_nd(const mat3<T>& A, const mat3<T>& B) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) s += A.e[i][j].r * B.e[i][j].r + A.e[i][j].i * B.e[i][j].i;
    return s;
}

// -P_TA(M) with P_TA(M) = (M - M^dag)/2 - tr((M - M^dag)/2)/3 * 1
// (proj_lie_su3, src/su3/utils.cpp:130-134), returned in 8 reals.
template <typename T>
__host__ __device__ inline alg8<T> neg_proj_ta(const mat3<T>& M) {
    alg8<T> z;
    const T h = T(0.5);
    T tr3 = (M.e[0][0].i + M.e[1][1].i + M.e[2][2].i) / T(3);
    z.a[0] = -(M.e[0][0].i - tr3);
    z.a[1] = -(M.e[1][1].i - tr3);
    z.a[2] = -h * (M.e[0][1].r - M.e[1][0].r);
    z.a[3] = -h * (M.e[0][1].i + M.e[1][0].i);
    z.a[4] = -h * (M.e[0][2].r - M.e[2][0].r);
    z.a[5] = -h * (M.e[0][2].i + M.e[2][0].i);
    z.a[6] = -h * (M.e[1][2].r - M.e[2][1].r);
    z.a[7] = -h * (M.e[1][2].i + M.e[2][1].i);
    return z;
}

template <typename T>
__host__ __device__ inline mat3<T> alg_to_mat(const alg8<T>& x) {
    mat3<T> m;
    T d2 = -x.a[0] - x.a[1];
    m.e[0][0] = {T(0), x.a[0]};
    m.e[1][1] = {T(0), x.a[1]};
    m.e[2][2] = {T(0), d2};
    m.e[0][1] = {x.a[2], x.a[3]};
    m.e[0][2] = {x.a[4], x.a[5]};
    m.e[1][2] = {x.a[6], x.a[7]};
    m.e[1][0] = {-x.a[2], x.a[3]};
    m.e[2][0] = {-x.a[4], x.a[5]};
    m.e[2][1] = {-x.a[6], x.a[7]};
    return m;
}

// Taylor coefficients c_n = 1/n!, n = 0..18, formed by repeated division
// exactly as exp_su3_luscher does (src/su3/utils.cpp:162-166).
struct horner_coeffs {
    double c[19];
};
__host__ __device__ constexpr horner_coeffs make_horner_coeffs() {
    horner_coeffs h{};
    h.c[0] = 1.0;
    for (int n = 1; n <= 18; n++) h.c[n] = h.c[n - 1] / static_cast<double>(n);
    return h;
}

// Cayley-Hamilton coefficients of exp(X) = q0 1 + q1 X + q2 X^2 for X in su(3)
// with t = -1/2 tr X^2 and d = Re(i det X), by the Horner recursion of
// exp_su3_luscher (src/su3/utils.cpp:168-186) from the term n = N, i.e. for
// the Taylor polynomial of degree N (the reference's N is 18):
//   q <- (c_n - i d q2, q0 - t q2, q1),  n = N-1..0.
// Evaluated in the type H (double in every path).
template <int N, typename H>
__host__ __device__ inline void horner_n(H t, H d, cpx<H>& q0, cpx<H>& q1, cpx<H>& q2) {
    static_assert(N >= 1 && N <= 18, "degree 1..18");
    constexpr horner_coeffs hc = make_horner_coeffs();
    cpx<H> a0 = {H(hc.c[N]), H(0)}, a1 = {H(0), H(0)}, a2 = {H(0), H(0)};
    // The first iterations multiply parts that are exactly zero (a1 and a2
    // start at 0; their imaginary parts fill in only from the fourth and
    // fifth iteration on).  The flags track those parts, so the products
    // with them are left out; for finite t and d that changes no value (c +
    // d*0 == c), and a fully unrolled loop folds the flags away.
    bool z0i = true, z1r = true, z1i = true, z2r = true, z2i = true;
#pragma unroll
    for (int n = N - 1; n >= 0; n--) {
        // c_n - i d q2
        const cpx<H> n0 = {z2i ? H(hc.c[n]) : H(hc.c[n]) + d * a2.i, z2r ? H(0) : -d * a2.r};
        // q0 - t q2
        const cpx<H> n1 = {z2r ? a0.r : a0.r - t * a2.r, z2i ? a0.i : a0.i - t * a2.i};
        const bool zn0i = z2r, zn1i = z0i && z2i;
        a2 = a1;
        z2r = z1r;
        z2i = z1i;
        a1 = n1;
        z1r = false;
        z1i = zn1i;
        a0 = n0;
        z0i = zn0i;
    }
    q0 = a0;
    q1 = a1;
    q2 = a2;
}
// Below this t the terms n = 11..18 of the 18-term series change exp(X) by
// less than 2^-56 (an eighth of an FP64 ulp of 1): |X| <= |X|_F = sqrt(2t) <
// 0.1415 and sum_{n>10} |X|^n/n! < |X|^11/11! e^|X| < 1.2e-17.  The
// exponents of a Wilson-flow step at eps = 0.02 stay below t = 2e-3, hot
// starts included.
#ifndef QCB_HORNER_SHORT
#define QCB_HORNER_SHORT 1
#endif
constexpr double kHornerShortT = 0.01;
// Where exp(X) is rounded to FP32 (the FP32-mixed flow), below this t the
// terms n = 8..18 change it by less than 2^-41: |X| < sqrt(2t) = 0.1 and
// sum_{n>7} |X|^n/n! < |X|^8/8! e^|X| < 2.8e-13, 1/2^17 of an FP32 ulp of 1.
// The coefficients are still evaluated in FP64, from the degree-7 polynomial.
constexpr double kHornerF32T = 0.005;

// The coefficients of the 18-term series, evaluated as the degree-10
// polynomial where that is equal to it within 2^-56, and for an FP32 result
// (F32OUT) as the degree-7 one where that is equal to it within 2^-41 (per
// lane, so the choice never depends on which sites share a wavefront).
template <typename H, bool F32OUT = false>
__host__ __device__ inline void horner_exp(H t, H d, cpx<H>& q0, cpx<H>& q1, cpx<H>& q2) {
    if (F32OUT && QCB_HORNER_SHORT && t < H(kHornerF32T))
        horner_n<7>(t, d, q0, q1, q2);
    else if (QCB_HORNER_SHORT && t < H(kHornerShortT))
        horner_n<10>(t, d, q0, q1, q2);
    else
        horner_n<18>(t, d, q0, q1, q2);
}

// exp(X) for X in su(3), following exp_su3_luscher (src/su3/utils.cpp:145-193).
// t and d use the closed forms for a traceless anti-hermitian X:
//   t = (d0^2 + d1^2 + d2^2)/2 + |p|^2 + |q|^2 + |r|^2
//   d = d0 d1 d2 - d0 |r|^2 - d1 |q|^2 - d2 |p|^2 + 2 Im(p r q*)
// Sets *warn when t > 1 (the reference prints a warning, utils.cpp:153-155).
// X^2 is hermitian, so only its diagonal and upper triangle are formed.
template <typename T, typename H>
__host__ __device__ inline mat3<T> exp_su3(const alg8<T>& x, bool* warn) {
    const T d0 = x.a[0], d1 = x.a[1], d2 = -x.a[0] - x.a[1];
    const cpx<T> p = {x.a[2], x.a[3]}, q = {x.a[4], x.a[5]}, r = {x.a[6], x.a[7]};
    const T pp = norm2(p), qq = norm2(q), rr = norm2(r);
    // X^2 (hermitian): diagonal real, upper triangle
    const T s00 = -(d0 * d0) - pp - qq;
    const T s11 = -pp - d1 * d1 - rr;
    const T s22 = -qq - rr - d2 * d2;
    // (X^2)01 = i(d0+d1) p - q r*
    const cpx<T> s01 = {-(d0 + d1) * p.i - (q.r * r.r + q.i * r.i),
                        (d0 + d1) * p.r - (q.i * r.r - q.r * r.i)};
    // (X^2)02 = i(d0+d2) q + p r
    const cpx<T> s02 = {-(d0 + d2) * q.i + (p.r * r.r - p.i * r.i),
                        (d0 + d2) * q.r + (p.r * r.i + p.i * r.r)};
    // (X^2)12 = -p* q + i(d1+d2) r
    const cpx<T> s12 = {-(p.r * q.r + p.i * q.i) - (d1 + d2) * r.i,
                        -(p.r * q.i - p.i * q.r) + (d1 + d2) * r.r};

    H hd0 = H(d0), hd1 = H(d1), hd2 = H(d2);
    H tt = H(0.5) * (hd0 * hd0 + hd1 * hd1 + hd2 * hd2) + H(pp) + H(qq) + H(rr);
    // Im(p r q*) = Im(p r) Re q - Re(p r) Im q
    H prr = H(p.r) * H(r.r) - H(p.i) * H(r.i);
    H pri = H(p.r) * H(r.i) + H(p.i) * H(r.r);
    H dd = hd0 * hd1 * hd2 - hd0 * H(rr) - hd1 * H(qq) - hd2 * H(pp) +
           H(2) * (pri * H(q.r) - prr * H(q.i));
    if (warn && tt > H(1)) *warn = true;
    cpx<H> h0, h1, h2;
    horner_exp<H, sizeof(T) == 4>(tt, dd, h0, h1, h2);
    const cpx<T> q0 = {T(h0.r), T(h0.i)}, q1 = {T(h1.r), T(h1.i)}, q2 = {T(h2.r), T(h2.i)};

    mat3<T> E;
    // diagonal: q0 + q1 (i d_k) + q2 s_kk
    E.e[0][0] = {q0.r - q1.i * d0 + q2.r * s00, q0.i + q1.r * d0 + q2.i * s00};
    E.e[1][1] = {q0.r - q1.i * d1 + q2.r * s11, q0.i + q1.r * d1 + q2.i * s11};
    E.e[2][2] = {q0.r - q1.i * d2 + q2.r * s22, q0.i + q1.r * d2 + q2.i * s22};
    // upper: q1 X_ij + q2 S_ij ; lower: q1 X_ji + q2 conj(S_ij), X_ji = -conj(X_ij)
    const cpx<T> mp = {-p.r, p.i}, mq = {-q.r, q.i}, mr = {-r.r, r.i};
    E.e[0][1] = madd(q1 * p, q2, s01);
    E.e[0][2] = madd(q1 * q, q2, s02);
    E.e[1][2] = madd(q1 * r, q2, s12);
    E.e[1][0] = madd(q1 * mp, q2, conj(s01));
    E.e[2][0] = madd(q1 * mq, q2, conj(s02));
    E.e[2][1] = madd(q1 * mr, q2, conj(s12));
    return E;
}

}  // namespace qcb
