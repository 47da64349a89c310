// flow.hip -- Wilson-flow kernels for gfx950, FP64 and FP32-mixed.
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
// The same code serves both storage forms (fields.hpp): the FP64 path does all
// arithmetic in FP64; the FP32-mixed path does the staples, force, RK3
// combination and the exp reconstruction in FP32 and keeps FP64 for the
// Cayley-Hamilton coefficients (t, d and the Horner recursion) and for the
// re-unitarisation of the stored rows.
//
// Work mapping: a workgroup is 64 consecutive sites x 4 directions; wavefront
// w handles direction mu = w, so every index computation inside a wavefront is
// uniform and the four wavefronts of one workgroup share neighbour links
// through the CU's L1.  With two or one threads per site (lattice::dpw) a
// wavefront updates two or four directions in turn and the workgroup takes
// two or four tiles (stage_directions, flow_body.hpp).  The workgroup ->
// site-tile order is XCD-aware (tile_of_block, lattice.hpp).
#include "flow_body.hpp"
#include "kernels.hpp"
#include "measure_body.hpp"
#include "reduce.hpp"

namespace qcb {

// Wavefronts per SIMD the stage kernel is compiled for: FP64 3 (its natural
// 150 VGPRs), FP32-mixed 6 (76).  Two ways to more FP64 wavefronts were
// measured slower on MI355X: staging the site's own links in LDS for the four
// direction wavefronts (the barrier costs more than the L1 hits it replaces),
// and keeping the staple sum in LDS (130 VGPRs at 3 waves, 128 + spills at 4).
// Four waves with loaded x-neighbours (128 VGPRs, at most 8 spilled, a
// scheduling barrier between the two staples of a plane) ran 7-12% slower
// than three waves with lane shuffles at 32^4 and 64^4.  Forming each
// plaquette once and handing -P_TA of it to the other direction's wavefront
// through LDS (U_mu(x) times the upper staple of plane (mu, nu) is the
// plaquette, the adjoint of wavefront nu's term) cut the VALU instructions
// by 7-8% but, with its barrier, was no faster than the lane shuffles in FP64
// and 5-13% slower in FP32-mixed.
template <prec P>
constexpr int kStageWaves = P == prec::f64 ? 3 : 6;

// FP64 stage kernels take the x-neighbour links of whole x-rows from lane
// shuffles (staple_sum, XROWS): fewer vector-memory instructions for a kernel
// whose loads keep the texture-address path about 60% busy.  The plain
// FP32-mixed stage kernels keep their loads (the shuffled words would cost
// them their sixth wavefront per SIMD); their default where whole x-rows fit
// a wavefront is the LDS-staged kernel (stage_lds.hip).
#ifndef QCB_XROW_SHUFFLES
#define QCB_XROW_SHUFFLES 1
#endif
template <prec P>
inline bool use_x_rows(const lattice& g) {
    return QCB_XROW_SHUFFLES && P == prec::f64 && x_rows_in_wave(g);
}

template <prec P, int STAGE, bool XROWS, int DPW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kStageWaves<P>))) void k_rk3_stage(
    const typename storage<P>::vec* __restrict__ Uin, typename storage<P>::vec* __restrict__ Uout,
    typename storage<P>::avec* __restrict__ acc, lattice g, double cz, double ca,
    int* __restrict__ warn) {
    const unsigned ntiles = stage_tiles(g);
    const tile_groups G = stage_groups(g);
    const unsigned grp = tile_of_block(blockIdx.x, G.n, g.map_mode, G.win, G.per_t);
    const unsigned tile = group_tile<DPW>(grp, threadIdx.y);
    if (tile >= ntiles) return;  // padding group, or past the last tile
    unsigned c[4];
    const unsigned s = tile_site(g, tile, threadIdx.x, c);
    if (s >= g.V) return;  // never a partial wavefront when XROWS (full tiles)
    nbr n = neighbours_at(g, c);
    if constexpr (XROWS) set_x_lanes(n, g, threadIdx.x);
    bool w = false;
    stage_directions<P, STAGE, XROWS, DPW>(threadIdx.y, Uin, Uout, acc, g.V, s, n, cz, ca, w);
    if (w) atomicOr(warn, 1);
}

template <prec P, int DPW>
static hipError_t stage_launch(int stage, const void* Uin, void* Uout, void* acc, const lattice& g,
                               double cz, double ca, int* warn, hipStream_t st) {
    using S = storage<P>;
    const auto* in = static_cast<const typename S::vec*>(Uin);
    auto* out = static_cast<typename S::vec*>(Uout);
    auto* a = static_cast<typename S::avec*>(acc);
    const dim3 block(kSitesPerBlock, 4);
    const tile_groups G = stage_groups(g);
    const dim3 grid(grid_for_tiles(G.n, g.map_mode, G.win, G.per_t));
    const bool xr = use_x_rows<P>(g);
    constexpr bool XR = P == prec::f64;  // only FP64 has x-row variants
    switch (stage * 2 + (xr ? 1 : 0)) {
        case 2: k_rk3_stage<P, 1, false, DPW><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 3: k_rk3_stage<P, 1, XR, DPW><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 4: k_rk3_stage<P, 2, false, DPW><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 5: k_rk3_stage<P, 2, XR, DPW><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 6: k_rk3_stage<P, 3, false, DPW><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        case 7: k_rk3_stage<P, 3, XR, DPW><<<grid, block, 0, st>>>(in, out, a, g, cz, ca, warn); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <prec P>
static hipError_t stage_launch(int stage, const void* Uin, void* Uout, void* acc, const lattice& g,
                               double cz, double ca, int* warn, hipStream_t st) {
    switch (g.dpw) {
        case 1: return stage_launch<P, 1>(stage, Uin, Uout, acc, g, cz, ca, warn, st);
        case 2: return stage_launch<P, 2>(stage, Uin, Uout, acc, g, cz, ca, warn, st);
        case 4: return stage_launch<P, 4>(stage, Uin, Uout, acc, g, cz, ca, warn, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rk3_stage(prec p, int stage, const void* Uin, void* Uout, void* acc,
                            const lattice& g, double cz, double ca, int* warn, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    return p == prec::f64 ? stage_launch<prec::f64>(stage, Uin, Uout, acc, g, cz, ca, warn, st)
                          : stage_launch<prec::f32m>(stage, Uin, Uout, acc, g, cz, ca, warn, st);
}

// ---- RK3 stage 1 fused with the measurement of its input -------------------
// topo_charge_flowed measures the field and then flows it, so every
// measurement but the last is followed by an RK3 stage 1 on the same field.
// That stage already forms the upper and lower staples of every link, which
// give three of the four leaves of each clover (clover_parts,
// measure_body.hpp); this kernel does the stage and the measurement in one
// pass, adding 8-10 SU(3) products per site and direction to the stage
// instead of a separate measurement.  Its per-tile partials are those of
// k_measure_clover (same per-site code, same workgroup reduction).  Stage
// tiles must be 64 consecutive sites (the default shape).
template <prec P, int MU, bool XROWS>
__device__ inline double stage1_measure_wave(const typename storage<P>::field& f,
                                             typename storage<P>::vec* __restrict__ Uout,
                                             typename storage<P>::avec* __restrict__ acc,
                                             unsigned s, const nbr& n, double cz, bool& warn,
                                             typename storage<P>::real* lds, unsigned lane) {
    clover_parts<typename storage<P>::field, MU> h{f, s, n, f.link(MU, s), lds, lane, 0.0, {}};
    stage_body<P, 1, MU, XROWS>(f, Uout, acc, s, n, cz, 0.0, warn, h);
    h.finish();
    return h.plaq;
}

// The fused measurement takes the x-neighbour links of whole x-rows from lane
// shuffles in both precisions: in FP32-mixed it is the kernel with the most
// loads per wavefront (the clover leaves' links on top of the stage's), and
// the shuffles ran it 7% faster at 64^4 (the same occupancy, 124 VGPRs).
#ifndef QCB_S1M_F32_XROWS
#define QCB_S1M_F32_XROWS 1
#endif

template <prec P, bool XROWS>
__global__ __launch_bounds__(256) void k_rk3_stage1_measure(
    const typename storage<P>::vec* __restrict__ Uin, typename storage<P>::vec* __restrict__ Uout,
    typename storage<P>::avec* __restrict__ acc, lattice g, double cz, int* __restrict__ warn,
    double* __restrict__ partials) {
    using FT = typename storage<P>::field;
    if constexpr (!kFusedMeasureBuilt) return;  // never launched (launcher refuses)
    __shared__ typename FT::real lds[kCloverSlots * 9 * 64];
    __shared__ double red[4 * 3];
    const unsigned ntiles = measure_blocks(g);
    const unsigned tile =
        tile_of_block(blockIdx.x, ntiles, g.map_mode, g.map_win, site_tiles_per_t(g));
    if (tile >= ntiles) return;  // a whole padding workgroup: no barrier is skipped
    const unsigned lane = threadIdx.x;
    const unsigned s = tile * kSitesPerBlock + lane;
    const bool valid = s < g.V;
    const FT f = storage<P>::view(Uin, g.V);
    nbr n = neighbours(g, valid ? s : 0u);
    if constexpr (XROWS) set_x_lanes(n, g, lane);  // full tiles: every lane is valid
    double v[3] = {0.0, 0.0, 0.0};
    bool w = false;
    if (valid) {
        switch (threadIdx.y) {
            case 0: v[0] = stage1_measure_wave<P, 0, XROWS>(f, Uout, acc, s, n, cz, w, lds, lane); break;
            case 1: v[0] = stage1_measure_wave<P, 1, XROWS>(f, Uout, acc, s, n, cz, w, lds, lane); break;
            case 2: v[0] = stage1_measure_wave<P, 2, XROWS>(f, Uout, acc, s, n, cz, w, lds, lane); break;
            default: v[0] = stage1_measure_wave<P, 3, XROWS>(f, Uout, acc, s, n, cz, w, lds, lane); break;
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
    if (w) atomicOr(warn, 1);
}

hipError_t launch_rk3_stage1_measure(prec p, const void* Uin, void* Uout, void* acc,
                                     const lattice& g, double cz, int* warn, double* partials,
                                     double* out, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (!kFusedMeasureBuilt || g.lg_tile[0] >= 0) return hipErrorInvalidValue;  // not fusable
    const unsigned nblk = measure_blocks(g);
    const dim3 grid(grid_for_tiles(nblk, g.map_mode, g.map_win, site_tiles_per_t(g)));
    const dim3 block(kSitesPerBlock, 4);
    const auto* i64 = static_cast<const double2*>(Uin);
    auto* o64 = static_cast<double2*>(Uout);
    auto* a64 = static_cast<double2*>(acc);
    const auto* i32 = static_cast<const float4*>(Uin);
    auto* o32 = static_cast<float4*>(Uout);
    auto* a32 = static_cast<float4*>(acc);
    if (p == prec::f64 && use_x_rows<prec::f64>(g))
        k_rk3_stage1_measure<prec::f64, true><<<grid, block, 0, st>>>(i64, o64, a64, g, cz, warn, partials);
    else if (p == prec::f64)
        k_rk3_stage1_measure<prec::f64, false><<<grid, block, 0, st>>>(i64, o64, a64, g, cz, warn, partials);
    else if (QCB_S1M_F32_XROWS && x_rows_in_wave(g))
        k_rk3_stage1_measure<prec::f32m, true><<<grid, block, 0, st>>>(i32, o32, a32, g, cz, warn, partials);
    else
        k_rk3_stage1_measure<prec::f32m, false><<<grid, block, 0, st>>>(i32, o32, a32, g, cz, warn, partials);
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_reduce_measure(partials, g, out, st);
}

// ---- unfused pieces (GradientFlow member functions) -------------------------

template <prec P, int MU>
__device__ inline void force_body(const typename storage<P>::field& f,
                                  typename storage<P>::avec* __restrict__ F, unsigned s,
                                  const nbr& n) {
    mat3<typename storage<P>::real> u;
    storage<P>::store_alg(F, f.V, MU, s, force_at<typename storage<P>::field, MU>(f, s, n, u));
}

// compute_force: src/flow/gradient_flow.cpp:17-37
template <prec P>
__global__ __launch_bounds__(256) void k_force(const typename storage<P>::vec* __restrict__ U,
                                               typename storage<P>::avec* __restrict__ F,
                                               lattice g) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= g.V) return;
    const nbr n = neighbours(g, s);
    const auto f = storage<P>::view(U, g.V);
    switch (threadIdx.y) {
        case 0: force_body<P, 0>(f, F, s, n); break;
        case 1: force_body<P, 1>(f, F, s, n); break;
        case 2: force_body<P, 2>(f, F, s, n); break;
        default: force_body<P, 3>(f, F, s, n); break;
    }
}

// compute_w1 / compute_w2 / compute_w3: src/flow/gradient_flow.cpp:40-79,102-121.
// One thread per link, in place.
template <prec P>
__global__ __launch_bounds__(256) void k_update(int which, typename storage<P>::vec* __restrict__ U,
                                                const typename storage<P>::avec* __restrict__ F0,
                                                const typename storage<P>::avec* __restrict__ F1,
                                                lattice g, double eps, int* __restrict__ warn) {
    using S = storage<P>;
    using T = typename S::real;
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= g.V) return;
    const int mu = threadIdx.y;
    const alg8<T> a0 = S::load_alg(F0, g.V, mu, s);
    alg8<T> x;
    if (which == 1) {
        const T c = (T)(0.25 * eps);
        for (int k = 0; k < 8; k++) x.a[k] = c * a0.a[k];
    } else {
        const alg8<T> a1 = S::load_alg(F1, g.V, mu, s);
        const T c1 = (T)((which == 2) ? (8.0 / 9.0) * eps : 0.75 * eps);
        const T c0 = (T)((which == 2) ? (17.0 / 36.0) * eps : eps);
        for (int k = 0; k < 8; k++) x.a[k] = c1 * a1.a[k] - c0 * a0.a[k];
    }
    bool w = false;
    const mat3<T> e = exp_su3<T, double>(x, &w);
    const mat3<T> u = S::view(U, g.V).link(mu, s);
    S::store_stage(U, g.V, mu, s, mul(e, u), which);
    if (w) atomicOr(warn, 1);
}

// replace_force_0: src/flow/gradient_flow.cpp:82-99.  F0 <- 8/9 F1 - 17/36 F0.
template <prec P>
__global__ __launch_bounds__(256) void k_replace(typename storage<P>::avec* __restrict__ F0,
                                                 const typename storage<P>::avec* __restrict__ F1,
                                                 lattice g) {
    using S = storage<P>;
    using T = typename S::real;
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= g.V) return;
    const int mu = threadIdx.y;
    const alg8<T> a0 = S::load_alg(F0, g.V, mu, s), a1 = S::load_alg(F1, g.V, mu, s);
    alg8<T> r;
    for (int k = 0; k < 8; k++) r.a[k] = (T)(8.0 / 9.0) * a1.a[k] - (T)(17.0 / 36.0) * a0.a[k];
    S::store_alg(F0, g.V, mu, s, r);
}

template <prec P>
static hipError_t unfused_launch(int which, void* U, void* F0, const void* F1, const lattice& g,
                                 double eps, int* warn, hipStream_t st) {
    using S = storage<P>;
    const dim3 block(kSitesPerBlock, 4);
    const dim3 grid((g.V + kSitesPerBlock - 1) / kSitesPerBlock);
    auto* u = static_cast<typename S::vec*>(U);
    auto* f0 = static_cast<typename S::avec*>(F0);
    const auto* f1 = static_cast<const typename S::avec*>(F1);
    if (which == 0)
        k_force<P><<<grid, block, 0, st>>>(u, f0, g);  // F0 is the output here
    else if (which <= 3)
        k_update<P><<<grid, block, 0, st>>>(which, u, f0, f1, g, eps, warn);
    else
        k_replace<P><<<grid, block, 0, st>>>(f0, f1, g);
    return hipGetLastError();
}

hipError_t launch_force(prec p, const void* U, void* F, const lattice& g, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    void* u = const_cast<void*>(U);
    return p == prec::f64 ? unfused_launch<prec::f64>(0, u, F, nullptr, g, 0.0, nullptr, st)
                          : unfused_launch<prec::f32m>(0, u, F, nullptr, g, 0.0, nullptr, st);
}
hipError_t launch_update(prec p, int which, void* U, const void* F0, const void* F1,
                         const lattice& g, double eps, int* warn, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (which < 1 || which > 3) return hipErrorInvalidValue;
    void* f0 = const_cast<void*>(F0);
    return p == prec::f64 ? unfused_launch<prec::f64>(which, U, f0, F1, g, eps, warn, st)
                          : unfused_launch<prec::f32m>(which, U, f0, F1, g, eps, warn, st);
}
hipError_t launch_replace_force0(prec p, void* F0, const void* F1, const lattice& g,
                                 hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    return p == prec::f64 ? unfused_launch<prec::f64>(4, nullptr, F0, F1, g, 0.0, nullptr, st)
                          : unfused_launch<prec::f32m>(4, nullptr, F0, F1, g, 0.0, nullptr, st);
}

}  // namespace qcb
