// fields.hpp -- device access to link fields and su(3) fields in HBM, for the
// two storage forms of the flow.
//
// field64  (plain path): 9 complex entries per link, FP64, structure of arrays
//          U[(mu*9 + k)*V + s] as double2 (lattice.hpp).
// field32c (FP32-mixed path): the first two rows only, FP32, as three float4
//          per link, U[(mu*3 + w)*V + s]:
//            w = 0: (U00, U01)   w = 1: (U02, U10)   w = 2: (U11, U12);
//          the third row is rebuilt on load as conj(row0 x row1), which is
//          exact for SU(3) (U^dag = adj U when det U = 1).  Uploads and the
//          flow's RK3 stage 3 re-unitarise the two rows in FP64 by row
//          Gram-Schmidt (the reference's proj_su3, src/su3/utils.cpp:194-214,
//          does the same by columns on full matrices; with two stored rows the
//          row form is the natural one), so the field stays in SU(3) to FP32
//          rounding; the intermediate fields of stages 1 and 2 keep the FP32
//          rows of exp(X) U, within a few FP32 roundings of SU(3).
// The FP64 flow output may hold rows 0-1 only (store_flow_link64,
// lattice.hpp); the FP32-mixed form never stores the third row.
// su(3) fields (force / RK3 accumulator) hold the 8 reals of alg8:
//          FP64 as 4 double2 per link, FP32 as 2 float4 per link.
#pragma once
#include <hip/hip_runtime.h>

#include "lattice.hpp"

// Products of links (staples, clover leaves) formed as SU(3) products: two
// rows and a cross product (mul_su3 and friends, su3.hpp).  0 = full products.
#ifndef QCB_SU3_PRODUCTS
#define QCB_SU3_PRODUCTS 1
#endif

// FP32-mixed flow: re-unitarise the stored rows after every RK3 stage (1) or
// after stage 3 only (0).
#ifndef QCB_F32M_PROJECT_EVERY_STAGE
#define QCB_F32M_PROJECT_EVERY_STAGE 0
#endif

namespace qcb {

enum class prec { f64 = 0, f32m = 1 };

// ---- FP32 two-row compressed field -------------------------------------------
__device__ inline float4 word_at(const float4* __restrict__ base, unsigned s) {
    return *word_ptr(base, s);
}

struct field32c {
    using real = float;
    using word = float4;
    static constexpr bool stream_rows = false;
    static constexpr int kWords = 3;  // two rows as three float4
    const float4* __restrict__ p;
    unsigned V;
    __device__ inline void load_words(int mu, unsigned s, word (&w)[kWords]) const {
#pragma unroll
        for (int k = 0; k < kWords; k++) w[k] = word_at(p + ((size_t)mu * 3 + k) * V, s);
    }
    __device__ static inline mat3<float> from_words(const word (&w)[kWords]) {
        const float4 a = w[0], c = w[1], d = w[2];
        mat3<float> m;
        m.e[0][0] = {a.x, a.y};
        m.e[0][1] = {a.z, a.w};
        m.e[0][2] = {c.x, c.y};
        m.e[1][0] = {c.z, c.w};
        m.e[1][1] = {d.x, d.y};
        m.e[1][2] = {d.z, d.w};
        m.e[2][0] = conj(m.e[0][1] * m.e[1][2] - m.e[0][2] * m.e[1][1]);
        m.e[2][1] = conj(m.e[0][2] * m.e[1][0] - m.e[0][0] * m.e[1][2]);
        m.e[2][2] = conj(m.e[0][0] * m.e[1][1] - m.e[0][1] * m.e[1][0]);
        return m;
    }
    __device__ inline mat3<float> link(int mu, unsigned s) const {
        word w[kWords];
        load_words(mu, s, w);
        return from_words(w);
    }
};

// One 16-byte store per lane: the word goes out as a vector, so the store
// never splits into narrower ones (which write their cache lines in parts).
typedef float float4_vec __attribute__((ext_vector_type(4)));
__device__ inline void put4(float4* p, float4 v) {
#if QCB_NT_STORES
    const float4_vec w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<float4_vec*>(p));
#else
    *p = v;
#endif
}

// 1/sqrt(n) in FP64 for the squared row norms of a link (n within FP32's
// normal range; about 1 for SU(3) rows): the hardware's FP32 reciprocal square
// root as the seed (relative error about 1e-7) and one FP64 Newton step
// y (3 - n y^2) / 2, relative error about 1e-14, far below the FP32 rounding
// of the stored rows.  A correctly rounded 1.0 / sqrt(n) costs an FP64 divide
// and square-root sequence per row.
__device__ inline double rsqrt_rows(double n) {
    const double y = (double)__builtin_amdgcn_rsqf((float)n);
    return y * (1.5 - 0.5 * n * y * y);
}

// FP64 Gram-Schmidt of two rows: r0 <- r0/|r0|, r1 <- (r1 - <r0,r1> r0)/|.|
__device__ inline void orthonormalise_rows(cpx<double> (&r0)[3], cpx<double> (&r1)[3]) {
    double n0 = norm2(r0[0]) + norm2(r0[1]) + norm2(r0[2]);
    const double i0 = rsqrt_rows(n0);
#pragma unroll
    for (int k = 0; k < 3; k++) r0[k] = i0 * r0[k];
    cpx<double> dot = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; k++) dot = madd_cl(dot, r0[k], r1[k]);
#pragma unroll
    for (int k = 0; k < 3; k++) r1[k] = r1[k] - dot * r0[k];
    const double n1 = norm2(r1[0]) + norm2(r1[1]) + norm2(r1[2]);
    const double i1 = rsqrt_rows(n1);
#pragma unroll
    for (int k = 0; k < 3; k++) r1[k] = i1 * r1[k];
}

__device__ inline void store_rows32c(float4* __restrict__ p, unsigned V, int mu, unsigned s,
                                     cpx<double> (&r0)[3], cpx<double> (&r1)[3]) {
    orthonormalise_rows(r0, r1);
    float4* b = p + (size_t)mu * 3 * V;
    put4(word_ptr(b, s),
         make_float4((float)r0[0].r, (float)r0[0].i, (float)r0[1].r, (float)r0[1].i));
    put4(word_ptr(b + V, s),
         make_float4((float)r0[2].r, (float)r0[2].i, (float)r1[0].r, (float)r1[0].i));
    put4(word_ptr(b + 2 * (size_t)V, s),
         make_float4((float)r1[1].r, (float)r1[1].i, (float)r1[2].r, (float)r1[2].i));
}

template <typename T>
__device__ inline void store_link32c(float4* __restrict__ p, unsigned V, int mu, unsigned s,
                                     const mat3<T>& m) {
    cpx<double> r0[3], r1[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r0[k] = {(double)m.e[0][k].r, (double)m.e[0][k].i};
        r1[k] = {(double)m.e[1][k].r, (double)m.e[1][k].i};
    }
    store_rows32c(p, V, mu, s, r0, r1);
}

// Rows 0-1 of an FP32 link stored as they are (no re-unitarisation).
__device__ inline void store_link32c_raw(float4* __restrict__ p, unsigned V, int mu, unsigned s,
                                         const mat3<float>& m) {
    float4* b = p + (size_t)mu * 3 * V;
    put4(word_ptr(b, s), make_float4(m.e[0][0].r, m.e[0][0].i, m.e[0][1].r, m.e[0][1].i));
    put4(word_ptr(b + V, s), make_float4(m.e[0][2].r, m.e[0][2].i, m.e[1][0].r, m.e[1][0].i));
    put4(word_ptr(b + 2 * (size_t)V, s),
         make_float4(m.e[1][1].r, m.e[1][1].i, m.e[1][2].r, m.e[1][2].i));
}

__device__ inline alg8<float> load_alg32(const float4* __restrict__ p, unsigned V, int mu,
                                         unsigned s) {
    const float4* b = p + (size_t)mu * 2 * V;
#if QCB_NT_ACC_LOADS
    const float* q0 = &word_ptr(b, s)->x;
    const float* q1 = &word_ptr(b + V, s)->x;
    return alg8<float>{{__builtin_nontemporal_load(q0), __builtin_nontemporal_load(q0 + 1),
                        __builtin_nontemporal_load(q0 + 2), __builtin_nontemporal_load(q0 + 3),
                        __builtin_nontemporal_load(q1), __builtin_nontemporal_load(q1 + 1),
                        __builtin_nontemporal_load(q1 + 2), __builtin_nontemporal_load(q1 + 3)}};
#else
    const float4 a = word_at(b, s), c = word_at(b + V, s);
    return alg8<float>{{a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w}};
#endif
}
__device__ inline void store_alg32(float4* __restrict__ p, unsigned V, int mu, unsigned s,
                                   const alg8<float>& a) {
    float4* b = p + (size_t)mu * 2 * V;
    put4(word_ptr(b, s), make_float4(a.a[0], a.a[1], a.a[2], a.a[3]));
    put4(word_ptr(b + V, s), make_float4(a.a[4], a.a[5], a.a[6], a.a[7]));
}

// ---- storage traits used by the generic kernels ------------------------------
template <prec P>
struct storage;

template <>
struct storage<prec::f64> {
    using real = double;
    using field = field64;
    using vec = double2;   // link-field word
    using avec = double2;  // su(3)-field word
    static constexpr int link_words = 9;
    static constexpr int alg_words = 4;
    __device__ static field view(const vec* p, unsigned V) { return field{p, V}; }
    // output of a flow update (two rows when the readers rebuild the third)
    __device__ static void store(vec* p, unsigned V, int mu, unsigned s, const mat3<double>& m) {
        store_flow_link64(p, V, mu, s, m);
    }
    // the output of RK3 stage `stage` (the same for every stage in FP64)
    __device__ static void store_stage(vec* p, unsigned V, int mu, unsigned s,
                                       const mat3<double>& m, int) {
        store_flow_link64(p, V, mu, s, m);
    }
    __device__ static alg8<double> load_alg(const avec* p, unsigned V, int mu, unsigned s) {
        return load_alg64(p, V, mu, s);
    }
    __device__ static void store_alg(avec* p, unsigned V, int mu, unsigned s,
                                     const alg8<double>& a) {
        store_alg64(p, V, mu, s, a);
    }
};

template <>
struct storage<prec::f32m> {
    using real = float;
    using field = field32c;
    using vec = float4;
    using avec = float4;
    static constexpr int link_words = 3;
    static constexpr int alg_words = 2;
    __device__ static field view(const vec* p, unsigned V) { return field{p, V}; }
    __device__ static void store(vec* p, unsigned V, int mu, unsigned s, const mat3<float>& m) {
        store_link32c(p, V, mu, s, m);
    }
    // the output of RK3 stage `stage`: re-unitarised in FP64 after stage 3
    // (and after every stage with QCB_F32M_PROJECT_EVERY_STAGE); the
    // intermediate fields of stages 1 and 2 keep the FP32 rows of exp(X) U.
    __device__ static void store_stage(vec* p, unsigned V, int mu, unsigned s,
                                       const mat3<float>& m, int stage) {
        if (QCB_F32M_PROJECT_EVERY_STAGE || stage == 3)
            store_link32c(p, V, mu, s, m);
        else
            store_link32c_raw(p, V, mu, s, m);
    }
    __device__ static alg8<float> load_alg(const avec* p, unsigned V, int mu, unsigned s) {
        return load_alg32(p, V, mu, s);
    }
    __device__ static void store_alg(avec* p, unsigned V, int mu, unsigned s,
                                     const alg8<float>& a) {
        store_alg32(p, V, mu, s, a);
    }
};

inline size_t link_field_bytes(prec p, unsigned V) {
    return (size_t)V * 4 * (p == prec::f64 ? 9 * sizeof(double2) : 3 * sizeof(float4));
}
inline size_t alg_field_bytes(prec p, unsigned V) {
    return (size_t)V * 4 * (p == prec::f64 ? 4 * sizeof(double2) : 2 * sizeof(float4));
}

}  // namespace qcb
