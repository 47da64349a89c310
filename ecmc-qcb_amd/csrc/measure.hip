// measure.hip -- fused clover / plaquette / Q / E measurement (FP64 and FP32-mixed).
//
// Replaces mean_plaquette_local (src/observables/observables_mpi.cpp:11-35),
// clover_site (:66-90), local_q_e_clover (:93-112) and topo_q_e_clover[_global]
// (:115-156) with one pass over the links.  The first clover leaf of plane
// (mu,nu) is exactly the plaquette product of mean_plaquette_local, so the
// plaquette comes for free.
//
// The FP32-mixed path forms the clovers in FP32 and accumulates every
// per-site contribution in FP64; the final reduction is Neumaier-compensated.
//
// Work mapping (clover): a workgroup is 64 consecutive sites x 4 directions.
// Wavefront mu forms, from the staples of U_mu(x), the clover leaves it owns in
// its three planes (clover_parts, measure_body.hpp) and leaves their
// anti-hermitian parts in LDS; after one barrier wavefronts 0-2 each take one
// plane pair, (0,1)(2,3), (0,2)(1,3) or (0,3)(1,2), and form one term of
// q = Re tr F01F23 - Re tr F02F13 + Re tr F03F12 and two of the six energy
// terms.  The fused RK3 stage 1 (flow.hip) runs the same per-site code.
//
// The plaquette-only measurement keeps one wavefront per plane pair.
#include "kernels.hpp"
#include "measure_body.hpp"
#include "reduce.hpp"

namespace qcb {

template <prec P>
__global__ __launch_bounds__(256) void k_measure_clover(const typename storage<P>::vec* __restrict__ U,
                                                        lattice g, double* __restrict__ partials) {
    using FT = typename storage<P>::field;
    using T = typename FT::real;
    __shared__ T lds[kCloverSlots * 9 * 64];
    __shared__ double red[4 * 3];
    const unsigned ntiles = measure_blocks(g);
    const unsigned tile =
        tile_of_block(blockIdx.x, ntiles, g.map_mode, g.map_win, site_tiles_per_t(g));
    if (tile >= ntiles) return;  // a whole padding workgroup: no barrier is skipped
    const unsigned lane = threadIdx.x;
    const unsigned s = tile * kSitesPerBlock + lane;
    const bool valid = s < g.V;
    double v[3] = {0.0, 0.0, 0.0};
    const FT f = storage<P>::view(U, g.V);
    const nbr n = neighbours(g, valid ? s : 0u);
    if (valid) {
        switch (threadIdx.y) {
            case 0: v[0] = clover_wave<FT, 0>(f, s, n, lds, lane); break;
            case 1: v[0] = clover_wave<FT, 1>(f, s, n, lds, lane); break;
            case 2: v[0] = clover_wave<FT, 2>(f, s, n, lds, lane); break;
            default: v[0] = clover_wave<FT, 3>(f, s, n, lds, lane); break;
        }
    }
    __syncthreads();
    if (valid) clover_pair_terms(lds, lane, (int)threadIdx.y, v[1], v[2]);
    block_sum<3>(v, red, 4);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        partials[3 * tile + 0] = v[0];
        partials[3 * tile + 1] = v[1];
        partials[3 * tile + 2] = v[2];
    }
}

template <prec P>
__global__ __launch_bounds__(192) void k_plaquette(const typename storage<P>::vec* __restrict__ U,
                                                   lattice g, double* __restrict__ partials) {
    using FT = typename storage<P>::field;
    __shared__ double lds[3 * 3];
    const unsigned ntiles = measure_blocks(g);
    const unsigned tile =
        tile_of_block(blockIdx.x, ntiles, g.map_mode, g.map_win, site_tiles_per_t(g));
    if (tile >= ntiles) return;  // a whole padding workgroup: no barrier is skipped
    const unsigned s = tile * kSitesPerBlock + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (s < g.V) {
        const nbr n = neighbours(g, s);
        const FT f = storage<P>::view(U, g.V);
        switch (threadIdx.y) {
            case 0:
                plaq_only<FT, 0, 1>(f, s, n, v[0]);
                plaq_only<FT, 2, 3>(f, s, n, v[0]);
                break;
            case 1:
                plaq_only<FT, 0, 2>(f, s, n, v[0]);
                plaq_only<FT, 1, 3>(f, s, n, v[0]);
                break;
            default:
                plaq_only<FT, 0, 3>(f, s, n, v[0]);
                plaq_only<FT, 1, 2>(f, s, n, v[0]);
                break;
        }
    }
    block_sum<3>(v, lds, 3);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        partials[3 * tile + 0] = v[0];
        partials[3 * tile + 1] = v[1];
        partials[3 * tile + 2] = v[2];
    }
}

// Ordered, Neumaier-compensated reduction of the per-workgroup partials and
// the reference's normalisations (tile_triples_sum, write_observables).
__global__ __launch_bounds__(kReduceThreads) void k_reduce_measure(
    const double* __restrict__ partials, unsigned nblk, double invV, double* __restrict__ out) {
    double v[3];
    tile_triples_sum(partials, nblk, v);
    if (threadIdx.x == 0) write_observables(v, invV, out);
}

template <prec P>
static hipError_t measure_launch(const void* U, const lattice& g, double* partials, double* out,
                                 bool clover, hipStream_t st) {
    const unsigned nblk = measure_blocks(g);
    const unsigned grid = grid_for_tiles(nblk, g.map_mode, g.map_win, site_tiles_per_t(g));
    const auto* u = static_cast<const typename storage<P>::vec*>(U);
    if (clover)
        k_measure_clover<P><<<grid, dim3(kSitesPerBlock, 4), 0, st>>>(u, g, partials);
    else
        k_plaquette<P><<<grid, dim3(kSitesPerBlock, 3), 0, st>>>(u, g, partials);
    k_reduce_measure<<<1, kReduceThreads, 0, st>>>(partials, nblk, 1.0 / (double)g.V, out);
    return hipGetLastError();
}

hipError_t launch_reduce_measure(const double* partials, const lattice& g, double* out,
                                 hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    k_reduce_measure<<<1, kReduceThreads, 0, st>>>(partials, measure_blocks(g), 1.0 / (double)g.V, out);
    return hipGetLastError();
}

hipError_t launch_measure(prec p, const void* U, const lattice& g, double* partials, double* out,
                          bool clover, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    return p == prec::f64 ? measure_launch<prec::f64>(U, g, partials, out, clover, st)
                          : measure_launch<prec::f32m>(U, g, partials, out, clover, st);
}

// ---- one site -----------------------------------------------------------------
// The link U_d(s) in FP64 whatever the storage form.
template <class F>
__device__ inline mat3<double> link64(const F& f, int d, unsigned s) {
    const auto m = f.link(d, s);
    mat3<double> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.e[i][j] = {(double)m.e[i][j].r, (double)m.e[i][j].i};
    return r;
}
template <typename T>
__device__ inline mat3<double> adj(const mat3<T>& a) {
    mat3<double> r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.e[i][j] = conj(a.e[j][i]);
    return r;
}

// F_munu(x) = (C - C^dag)/8 of clover_site (observables_mpi.cpp:66-90): the four
// leaves in the reference's order and left-to-right association, each
// starting and ending at x, counter-clockwise in the (mu, nu) plane; zero for
// mu == nu (where the reference warns).
template <class F>
__device__ inline mat3<double> clover_at(const F& f, const lattice& g, unsigned x, int mu, int nu) {
    mat3<double> Fm = zero3<double>();
    if (mu == nu) return Fm;
    const nbr nx = neighbours(g, x);
    const unsigned xpmu = x + nx.up[mu], xpnu = x + nx.up[nu];
    const unsigned xmmu = x + nx.dn[mu], xmnu = x + nx.dn[nu];
    const nbr nm = neighbours(g, xmmu), np = neighbours(g, xpmu);
    const unsigned xmmupnu = xmmu + nm.up[nu], xmmumnu = xmmu + nm.dn[nu];
    const unsigned xpmumnu = xpmu + np.dn[nu];
    mat3<double> c = mul(mul(mul(link64(f, mu, x), link64(f, nu, xpmu)), adj(link64(f, mu, xpnu))),
                         adj(link64(f, nu, x)));
    add_to(c, mul(mul(mul(link64(f, nu, x), adj(link64(f, mu, xmmupnu))), adj(link64(f, nu, xmmu))),
                  link64(f, mu, xmmu)));
    add_to(c, mul(mul(mul(adj(link64(f, mu, xmmu)), adj(link64(f, nu, xmmumnu))),
                      link64(f, mu, xmmumnu)),
                  link64(f, nu, xmnu)));
    add_to(c, mul(mul(mul(adj(link64(f, nu, xmnu)), link64(f, mu, xmnu)), link64(f, nu, xpmumnu)),
                  adj(link64(f, mu, x))));
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Fm.e[i][j] = 0.125 * (c.e[i][j] - conj(c.e[j][i]));
    return Fm;
}

// Re tr(A B)
__device__ inline double re_tr_mul(const mat3<double>& A, const mat3<double>& B) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) s += A.e[i][k].r * B.e[k][i].r - A.e[i][k].i * B.e[k][i].i;
    return s;
}
__device__ inline double sq_norm(const mat3<double>& A) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) s += norm2(A.e[i][j]);
    return s;
}

// One thread: out[0..17] = F_munu at site x (row-major, re/im).
template <prec P>
__global__ void k_clover_site(const typename storage<P>::vec* __restrict__ U, lattice g,
                              unsigned x, int mu, int nu, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const mat3<double> F = clover_at(storage<P>::view(U, g.V), g, x, mu, nu);
    for (int k = 0; k < 9; k++) {
        out[2 * k] = F.e[k / 3][k % 3].r;
        out[2 * k + 1] = F.e[k / 3][k % 3].i;
    }
}

// One thread: local_q_e_clover (observables_mpi.cpp:93-112) at site x,
// out = {q(x) = (Re tr F01F23 - Re tr F02F13 + Re tr F03F12) / 4pi^2,
//        e(x) = 1/2 sum_{mu<nu} |F_munu|^2}.
template <prec P>
__global__ void k_local_q_e(const typename storage<P>::vec* __restrict__ U, lattice g,
                            unsigned x, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const auto f = storage<P>::view(U, g.V);
    const mat3<double> F01 = clover_at(f, g, x, 0, 1), F02 = clover_at(f, g, x, 0, 2);
    const mat3<double> F03 = clover_at(f, g, x, 0, 3), F12 = clover_at(f, g, x, 1, 2);
    const mat3<double> F13 = clover_at(f, g, x, 1, 3), F23 = clover_at(f, g, x, 2, 3);
    const double q = re_tr_mul(F01, F23) - re_tr_mul(F02, F13) + re_tr_mul(F03, F12);
    const double e = 0.5 * (sq_norm(F01) + sq_norm(F02) + sq_norm(F03) + sq_norm(F12) +
                            sq_norm(F13) + sq_norm(F23));
    out[0] = q * (1.0 / (4.0 * M_PI * M_PI));
    out[1] = e;
}

hipError_t launch_clover_site(prec p, const void* U, const lattice& g, unsigned site, int mu,
                              int nu, double* out, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (site >= g.V || mu < 0 || mu > 3 || nu < 0 || nu > 3) return hipErrorInvalidValue;
    if (p == prec::f64)
        k_clover_site<prec::f64><<<1, 64, 0, st>>>(static_cast<const double2*>(U), g, site, mu, nu, out);
    else
        k_clover_site<prec::f32m><<<1, 64, 0, st>>>(static_cast<const float4*>(U), g, site, mu, nu, out);
    return hipGetLastError();
}

hipError_t launch_local_q_e(prec p, const void* U, const lattice& g, unsigned site, double* out,
                            hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (site >= g.V) return hipErrorInvalidValue;
    if (p == prec::f64)
        k_local_q_e<prec::f64><<<1, 64, 0, st>>>(static_cast<const double2*>(U), g, site, out);
    else
        k_local_q_e<prec::f32m><<<1, 64, 0, st>>>(static_cast<const float4*>(U), g, site, out);
    return hipGetLastError();
}

// ---- unitarity check --------------------------------------------------------
template <prec P>
__global__ __launch_bounds__(256) void k_unitarity(const typename storage<P>::vec* __restrict__ U,
                                                   lattice g, double* __restrict__ partials) {
    __shared__ double lds[4 * 2];
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    double um = 0.0, dm = 0.0;
    if (s < g.V) {
        const mat3<typename storage<P>::real> uu = storage<P>::view(U, g.V).link(threadIdx.y, s);
        mat3<double> u;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) u.e[i][j] = {(double)uu.e[i][j].r, (double)uu.e[i][j].i};
        const mat3<double> p = mul_nd(u, u);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const double re = p.e[i][j].r - (i == j ? 1.0 : 0.0);
                um = fmax(um, fmax(fabs(re), fabs(p.e[i][j].i)));
            }
        const cpx<double> det =
            u.e[0][0] * (u.e[1][1] * u.e[2][2] - u.e[1][2] * u.e[2][1]) -
            u.e[0][1] * (u.e[1][0] * u.e[2][2] - u.e[1][2] * u.e[2][0]) +
            u.e[0][2] * (u.e[1][0] * u.e[2][1] - u.e[1][1] * u.e[2][0]);
        dm = sqrt((det.r - 1.0) * (det.r - 1.0) + det.i * det.i);
        if (!(um == um)) um = INFINITY;  // NaN -> inf so that max() keeps it
        if (!(dm == dm)) dm = INFINITY;
    }
    um = wave_max(um);
    dm = wave_max(dm);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) {
        lds[(tid >> 6) * 2] = um;
        lds[(tid >> 6) * 2 + 1] = dm;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 4; w++) {
            um = fmax(um, lds[2 * w]);
            dm = fmax(dm, lds[2 * w + 1]);
        }
        partials[2 * blockIdx.x] = um;
        partials[2 * blockIdx.x + 1] = dm;
    }
}

__global__ __launch_bounds__(256) void k_reduce_max2(const double* __restrict__ partials,
                                                     unsigned nblk, double* __restrict__ out) {
    __shared__ double lds[4 * 2];
    double a = 0.0, b = 0.0;
    for (unsigned k = threadIdx.x; k < nblk; k += blockDim.x) {
        a = fmax(a, partials[2 * k]);
        b = fmax(b, partials[2 * k + 1]);
    }
    a = wave_max(a);
    b = wave_max(b);
    if ((threadIdx.x & 63) == 0) {
        lds[(threadIdx.x >> 6) * 2] = a;
        lds[(threadIdx.x >> 6) * 2 + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            a = fmax(a, lds[2 * w]);
            b = fmax(b, lds[2 * w + 1]);
        }
        out[0] = a;
        out[1] = b;
    }
}

hipError_t launch_unitarity(prec p, const void* U, const lattice& g, double* partials,
                            double* out, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const unsigned nblk = measure_blocks(g);
    if (p == prec::f64)
        k_unitarity<prec::f64><<<nblk, dim3(kSitesPerBlock, 4), 0, st>>>(
            static_cast<const double2*>(U), g, partials);
    else
        k_unitarity<prec::f32m><<<nblk, dim3(kSitesPerBlock, 4), 0, st>>>(
            static_cast<const float4*>(U), g, partials);
    k_reduce_max2<<<1, 256, 0, st>>>(partials, nblk, out);
    return hipGetLastError();
}

}  // namespace qcb
