// flow_f64.hip -- FP64 Wilson-flow kernels for gfx950.
//
// The reference runs one RK3 step as seven OpenMP passes plus three halo
// exchanges (GradientFlow::rk3_step, src/flow/gradient_flow.cpp:124-135).
// Here each RK3 stage is ONE kernel: for every link (x, mu) it forms the staple
// sum (GaugeField::compute_staple, src/gauge/GaugeField.cpp:74-97), the force
// Z = P_TA(-U Sigma) (compute_force, gradient_flow.cpp:17-37), the RK3
// combination with an 8-real su(3) accumulator (compute_w1/w2/w3 and
// replace_force_0, gradient_flow.cpp:40-121), the Luscher exponential
// (exp_su3_luscher, src/su3/utils.cpp:145-193) and the update exp(X) U.
// Neighbours still need the old links, so the new links go to a second
// (ping-pong) buffer.
//
// Work mapping: a workgroup is 64 consecutive sites x 4 directions; wavefront
// w handles direction mu = w, so every index computation inside a wavefront is
// uniform and the four wavefronts of one workgroup share neighbour links
// through the CU's L1.
#include "kernels.hpp"

namespace qcb {

// Staple sum of U_mu(x):
//   sum_{nu != mu} U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag
//                + (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
// One operand of every product is streamed from memory a row at a time, so
// only the accumulator, one product and one full operand are live at once.
template <int MU>
__device__ inline mat3<double> staple_sum64(const field64& f, unsigned s, const nbr& n) {
    mat3<double> st = zero3<double>();
    const unsigned xmu = s + n.up[MU];
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
        if (nu == MU) continue;
        const unsigned xnu = s + n.up[nu];
        const unsigned xmnu = s + n.dn[nu];
        const unsigned xmunu = xmu + n.dn[nu];
        {  // T = U_nu(x) U_mu(x+nu); st += U_nu(x+mu) T^dag
            const mat3<double> B = f.link(MU, xnu);
            mat3<double> T;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                cpx<double> a[3];
                f.row(nu, s, i, a);
#pragma unroll
                for (int j = 0; j < 3; j++)
                    T.e[i][j] = madd(madd(a[0] * B.e[0][j], a[1], B.e[1][j]), a[2], B.e[2][j]);
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                cpx<double> c[3];
                f.row(nu, xmu, i, c);
#pragma unroll
                for (int j = 0; j < 3; j++)
                    st.e[i][j] = madd_c(madd_c(madd_c(st.e[i][j], c[0], T.e[j][0]), c[1], T.e[j][1]),
                                        c[2], T.e[j][2]);
            }
        }
        {  // T = U_mu(x-nu) U_nu(x+mu-nu); st += T^dag U_nu(x-nu)
            const mat3<double> B = f.link(nu, xmunu);
            mat3<double> T;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                cpx<double> a[3];
                f.row(MU, xmnu, i, a);
#pragma unroll
                for (int j = 0; j < 3; j++)
                    T.e[i][j] = madd(madd(a[0] * B.e[0][j], a[1], B.e[1][j]), a[2], B.e[2][j]);
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                cpx<double> v[3];
                f.row(nu, xmnu, k, v);  // row k of U_nu(x-nu)
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) st.e[i][j] = madd_cl(st.e[i][j], T.e[k][i], v[j]);
            }
        }
    }
    return st;
}

template <int MU>
__device__ inline alg8<double> force64(const field64& f, unsigned s, const nbr& n,
                                       mat3<double>& u) {
    const mat3<double> st = staple_sum64<MU>(f, s, n);
    u = f.link(MU, s);
    return neg_proj_ta(mul(u, st));
}

template <int STAGE, int MU>
__device__ inline void stage_body64(const field64& f, double2* __restrict__ Uout,
                                    double2* __restrict__ acc, unsigned s, const nbr& n,
                                    double cz, double ca, bool& warn) {
    mat3<double> u;
    const alg8<double> z = force64<MU>(f, s, n, u);
    alg8<double> x;
    if constexpr (STAGE == 1) {
#pragma unroll
        for (int k = 0; k < 8; k++) x.a[k] = cz * z.a[k];
        store_alg64(acc, f.V, MU, s, z);
    } else {
        alg8<double> a = load_alg64(acc, f.V, MU, s);
#pragma unroll
        for (int k = 0; k < 8; k++) x.a[k] = cz * z.a[k] - ca * a.a[k];
        if constexpr (STAGE == 2) {
#pragma unroll
            for (int k = 0; k < 8; k++) a.a[k] = (8.0 / 9.0) * z.a[k] - (17.0 / 36.0) * a.a[k];
            store_alg64(acc, f.V, MU, s, a);
        }
    }
    const mat3<double> e = exp_su3<double, double>(x, &warn);
    store_link64(Uout, f.V, MU, s, mul(e, u));
}

#ifndef QCB_STAGE_WAVES_PER_EU
#define QCB_STAGE_WAVES_PER_EU 1
#endif

template <int STAGE>
__global__ __launch_bounds__(256, QCB_STAGE_WAVES_PER_EU) void k_rk3_stage64(
    const double2* __restrict__ Uin, double2* __restrict__ Uout, double2* __restrict__ acc,
    lattice g, double cz, double ca, int* __restrict__ warn) {
    const unsigned ntiles = (g.V + kSitesPerBlock - 1) / kSitesPerBlock;
    const unsigned tile = tile_of_block(blockIdx.x, ntiles, g.map_mode, g.map_win);
    const unsigned s = tile * kSitesPerBlock + threadIdx.x;
    if (tile >= ntiles || s >= g.V) return;
    const nbr n = neighbours(g, s);
    const field64 f{Uin, g.V};
    bool w = false;
    switch (threadIdx.y) {
        case 0: stage_body64<STAGE, 0>(f, Uout, acc, s, n, cz, ca, w); break;
        case 1: stage_body64<STAGE, 1>(f, Uout, acc, s, n, cz, ca, w); break;
        case 2: stage_body64<STAGE, 2>(f, Uout, acc, s, n, cz, ca, w); break;
        default: stage_body64<STAGE, 3>(f, Uout, acc, s, n, cz, ca, w); break;
    }
    if (w) atomicOr(warn, 1);
}

hipError_t launch_rk3_stage_f64(int stage, const double2* Uin, double2* Uout, double2* acc,
                                const lattice& g, double cz, double ca, int* warn,
                                hipStream_t st) {
    const dim3 block(kSitesPerBlock, 4);
    const dim3 grid(grid_for_tiles((g.V + kSitesPerBlock - 1) / kSitesPerBlock, g.map_mode,
                                   g.map_win));
    switch (stage) {
        case 1: k_rk3_stage64<1><<<grid, block, 0, st>>>(Uin, Uout, acc, g, cz, ca, warn); break;
        case 2: k_rk3_stage64<2><<<grid, block, 0, st>>>(Uin, Uout, acc, g, cz, ca, warn); break;
        case 3: k_rk3_stage64<3><<<grid, block, 0, st>>>(Uin, Uout, acc, g, cz, ca, warn); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- unfused pieces (GradientFlow member functions) -------------------------

template <int MU>
__device__ inline void force_body64(const field64& f, double2* __restrict__ F, unsigned s,
                                    const nbr& n) {
    mat3<double> u;
    store_alg64(F, f.V, MU, s, force64<MU>(f, s, n, u));
}

// compute_force: src/flow/gradient_flow.cpp:17-37
__global__ __launch_bounds__(256) void k_force64(const double2* __restrict__ U,
                                                 double2* __restrict__ F, lattice g) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= g.V) return;
    const nbr n = neighbours(g, s);
    const field64 f{U, g.V};
    switch (threadIdx.y) {
        case 0: force_body64<0>(f, F, s, n); break;
        case 1: force_body64<1>(f, F, s, n); break;
        case 2: force_body64<2>(f, F, s, n); break;
        default: force_body64<3>(f, F, s, n); break;
    }
}

hipError_t launch_force_f64(const double2* U, double2* F, const lattice& g, hipStream_t st) {
    const dim3 block(kSitesPerBlock, 4);
    const dim3 grid((g.V + kSitesPerBlock - 1) / kSitesPerBlock);
    k_force64<<<grid, block, 0, st>>>(U, F, g);
    return hipGetLastError();
}

// compute_w1 / compute_w2 / compute_w3: src/flow/gradient_flow.cpp:40-79,102-121.
// One thread per link, in place.
__global__ __launch_bounds__(256) void k_update64(int which, double2* __restrict__ U,
                                                  const double2* __restrict__ F0,
                                                  const double2* __restrict__ F1, lattice g,
                                                  double eps, int* __restrict__ warn) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= g.V) return;
    const int mu = threadIdx.y;
    const alg8<double> a0 = load_alg64(F0, g.V, mu, s);
    alg8<double> x;
    if (which == 1) {
        for (int k = 0; k < 8; k++) x.a[k] = (0.25 * eps) * a0.a[k];
    } else {
        const alg8<double> a1 = load_alg64(F1, g.V, mu, s);
        const double c1 = (which == 2) ? (8.0 / 9.0) * eps : 0.75 * eps;
        const double c0 = (which == 2) ? (17.0 / 36.0) * eps : eps;
        for (int k = 0; k < 8; k++) x.a[k] = c1 * a1.a[k] - c0 * a0.a[k];
    }
    bool w = false;
    const mat3<double> e = exp_su3<double, double>(x, &w);
    const field64 f{U, g.V};
    const mat3<double> u = f.link(mu, s);
    store_link64(U, g.V, mu, s, mul(e, u));
    if (w) atomicOr(warn, 1);
}

hipError_t launch_update_f64(int which, double2* U, const double2* F0, const double2* F1,
                             const lattice& g, double eps, int* warn, hipStream_t st) {
    if (which < 1 || which > 3) return hipErrorInvalidValue;
    const dim3 block(kSitesPerBlock, 4);
    const dim3 grid((g.V + kSitesPerBlock - 1) / kSitesPerBlock);
    k_update64<<<grid, block, 0, st>>>(which, U, F0, F1, g, eps, warn);
    return hipGetLastError();
}

// replace_force_0: src/flow/gradient_flow.cpp:82-99.  F0 <- 8/9 F1 - 17/36 F0.
__global__ __launch_bounds__(256) void k_replace64(double2* __restrict__ F0,
                                                   const double2* __restrict__ F1, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 a = F0[i], b = F1[i];
    F0[i] = make_double2((8.0 / 9.0) * b.x - (17.0 / 36.0) * a.x,
                         (8.0 / 9.0) * b.y - (17.0 / 36.0) * a.y);
}

hipError_t launch_replace_force0_f64(double2* F0, const double2* F1, const lattice& g,
                                     hipStream_t st) {
    const size_t n = (size_t)g.V * 16;
    k_replace64<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(F0, F1, n);
    return hipGetLastError();
}

}  // namespace qcb
