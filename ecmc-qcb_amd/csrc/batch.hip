// batch.hip -- persistent ensemble kernels: one launch per RK3 stage (and one
// measurement + one reduction per measurement) for a whole ensemble of
// lattices of uneven volumes.
//
// The persistent grid is sized to the resident capacity of the GPU; workgroup b
// visits virtual slots v = b, b + grid, b + 2 grid, ... and each slot maps to a
// tile of the flat (configuration, tile) list through the XCD-aware order of
// tile_of_block.  The per-link and per-site work is exactly that of the
// single-lattice kernels (flow_body.hpp, measure_body.hpp).
#include "batch.hpp"
#include "flow_body.hpp"
#include "kernels.hpp"
#include "measure_body.hpp"
#include "reduce.hpp"

namespace qcb {

template <prec P, int STAGE>
__global__ __launch_bounds__(256) void k_batch_stage(const typename storage<P>::vec* __restrict__ Uin,
                                                     typename storage<P>::vec* __restrict__ Uout,
                                                     typename storage<P>::avec* __restrict__ acc,
                                                     const cfg_desc* __restrict__ cfgs,
                                                     const unsigned* __restrict__ tile_cfg,
                                                     unsigned ntiles, int map_mode, unsigned map_win,
                                                     double cz, double ca, int* __restrict__ warn) {
    const unsigned nslots = grid_for_tiles(ntiles, map_mode, map_win);
    bool w = false;
    for (unsigned v = blockIdx.x; v < nslots; v += gridDim.x) {
        const unsigned tile = tile_of_block(v, ntiles, map_mode, map_win);
        if (tile >= ntiles) continue;
        const cfg_desc& c = cfgs[tile_cfg[tile]];
        unsigned co[4];
        const unsigned s = tile_site(c.g, tile - c.tile0, threadIdx.x, co);
        if (s >= c.g.V) continue;
        const nbr n = neighbours_at(c.g, co);
        const auto f = storage<P>::view(Uin + c.u_off, c.g.V);
        auto* out = Uout + c.u_off;
        auto* a = acc + c.a_off;
        switch (threadIdx.y) {
            case 0: stage_body<P, STAGE, 0>(f, out, a, s, n, cz, ca, w); break;
            case 1: stage_body<P, STAGE, 1>(f, out, a, s, n, cz, ca, w); break;
            case 2: stage_body<P, STAGE, 2>(f, out, a, s, n, cz, ca, w); break;
            default: stage_body<P, STAGE, 3>(f, out, a, s, n, cz, ca, w); break;
        }
    }
    if (w) atomicOr(warn, 1);
}

template <prec P>
static hipError_t batch_stage_launch(int stage, const void* Uin, void* Uout, void* acc,
                                     const cfg_desc* cfgs, const unsigned* tile_cfg,
                                     unsigned ntiles, const batch_launch& bl, double cz, double ca,
                                     int* warn, hipStream_t st) {
    using S = storage<P>;
    const auto* in = static_cast<const typename S::vec*>(Uin);
    auto* out = static_cast<typename S::vec*>(Uout);
    auto* a = static_cast<typename S::avec*>(acc);
    const dim3 block(kSitesPerBlock, 4);
    const int mm = bl.map_mode;
    const unsigned mw = bl.map_win;
    switch (stage) {
        case 1:
            k_batch_stage<P, 1><<<bl.grid, block, 0, st>>>(in, out, a, cfgs, tile_cfg, ntiles, mm,
                                                           mw, cz, ca, warn);
            break;
        case 2:
            k_batch_stage<P, 2><<<bl.grid, block, 0, st>>>(in, out, a, cfgs, tile_cfg, ntiles, mm,
                                                           mw, cz, ca, warn);
            break;
        case 3:
            k_batch_stage<P, 3><<<bl.grid, block, 0, st>>>(in, out, a, cfgs, tile_cfg, ntiles, mm,
                                                           mw, cz, ca, warn);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_stage(prec p, int stage, const void* Uin, void* Uout, void* acc,
                              const cfg_desc* cfgs, const unsigned* tile_cfg, unsigned ntiles,
                              const batch_launch& bl, double cz, double ca, int* warn,
                              hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    return p == prec::f64
               ? batch_stage_launch<prec::f64>(stage, Uin, Uout, acc, cfgs, tile_cfg, ntiles, bl,
                                               cz, ca, warn, st)
               : batch_stage_launch<prec::f32m>(stage, Uin, Uout, acc, cfgs, tile_cfg, ntiles, bl,
                                                cz, ca, warn, st);
}

// ---- measurement ---------------------------------------------------------------
// Same per-site code and workgroup reduction as k_measure_clover (measure.hip):
// wavefront mu forms its clover leaves, LDS joins the two halves of each plane.
template <prec P>
__global__ __launch_bounds__(256) void k_batch_measure(const typename storage<P>::vec* __restrict__ U,
                                                       const cfg_desc* __restrict__ cfgs,
                                                       const unsigned* __restrict__ tile_cfg,
                                                       unsigned ntiles, int map_mode,
                                                       unsigned map_win,
                                                       double* __restrict__ partials) {
    using FT = typename storage<P>::field;
    __shared__ typename FT::real lds[kCloverSlots * 9 * 64];
    __shared__ double red[4 * 3];
    const unsigned nslots = grid_for_tiles(ntiles, map_mode, map_win);
    const unsigned lane = threadIdx.x;
    for (unsigned v0 = blockIdx.x; v0 < nslots; v0 += gridDim.x) {  // uniform per workgroup
        const unsigned tile = tile_of_block(v0, ntiles, map_mode, map_win);
        if (tile >= ntiles) continue;
        const cfg_desc& c = cfgs[tile_cfg[tile]];
        const unsigned s = (tile - c.tile0) * kSitesPerBlock + lane;
        const bool valid = s < c.g.V;
        const FT f = storage<P>::view(U + c.u_off, c.g.V);
        const nbr n = neighbours(c.g, valid ? s : 0u);
        double r[3] = {0.0, 0.0, 0.0};
        if (valid) {
            switch (threadIdx.y) {
                case 0: r[0] = clover_wave<FT, 0>(f, s, n, lds, lane); break;
                case 1: r[0] = clover_wave<FT, 1>(f, s, n, lds, lane); break;
                case 2: r[0] = clover_wave<FT, 2>(f, s, n, lds, lane); break;
                default: r[0] = clover_wave<FT, 3>(f, s, n, lds, lane); break;
            }
        }
        __syncthreads();
        if (valid) clover_pair_terms(lds, lane, (int)threadIdx.y, r[1], r[2]);
        block_sum<3>(r, red, 4);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            partials[3 * tile + 0] = r[0];
            partials[3 * tile + 1] = r[1];
            partials[3 * tile + 2] = r[2];
        }
        __syncthreads();  // lds and red are reused by the next tile
    }
}

// RK3 stage 1 fused with the measurement of its input, for every member
// (k_rk3_stage1_measure, flow.hip); members must use 64-consecutive-site tiles.
template <prec P, int MU>
__device__ inline double batch_stage1_measure_wave(const typename storage<P>::field& f,
                                                   typename storage<P>::vec* __restrict__ out,
                                                   typename storage<P>::avec* __restrict__ a,
                                                   unsigned s, const nbr& n, double cz, bool& w,
                                                   typename storage<P>::real* lds, unsigned lane) {
    clover_parts<typename storage<P>::field, MU> h{f, s, n, f.link(MU, s), lds, lane, 0.0, {}};
    stage_body<P, 1, MU>(f, out, a, s, n, cz, 0.0, w, h);
    h.finish();
    return h.plaq;
}

template <prec P>
__global__ __launch_bounds__(256) void k_batch_stage1_measure(
    const typename storage<P>::vec* __restrict__ Uin, typename storage<P>::vec* __restrict__ Uout,
    typename storage<P>::avec* __restrict__ acc, const cfg_desc* __restrict__ cfgs,
    const unsigned* __restrict__ tile_cfg, unsigned ntiles, int map_mode, unsigned map_win,
    double cz, int* __restrict__ warn, double* __restrict__ partials) {
    using FT = typename storage<P>::field;
    __shared__ typename FT::real lds[kCloverSlots * 9 * 64];
    __shared__ double red[4 * 3];
    const unsigned nslots = grid_for_tiles(ntiles, map_mode, map_win);
    const unsigned lane = threadIdx.x;
    bool w = false;
    for (unsigned v0 = blockIdx.x; v0 < nslots; v0 += gridDim.x) {  // uniform per workgroup
        const unsigned tile = tile_of_block(v0, ntiles, map_mode, map_win);
        if (tile >= ntiles) continue;
        const cfg_desc& c = cfgs[tile_cfg[tile]];
        const unsigned s = (tile - c.tile0) * kSitesPerBlock + lane;
        const bool valid = s < c.g.V;
        const FT f = storage<P>::view(Uin + c.u_off, c.g.V);
        const nbr n = neighbours(c.g, valid ? s : 0u);
        auto* out = Uout + c.u_off;
        auto* a = acc + c.a_off;
        double r[3] = {0.0, 0.0, 0.0};
        if (valid) {
            switch (threadIdx.y) {
                case 0: r[0] = batch_stage1_measure_wave<P, 0>(f, out, a, s, n, cz, w, lds, lane); break;
                case 1: r[0] = batch_stage1_measure_wave<P, 1>(f, out, a, s, n, cz, w, lds, lane); break;
                case 2: r[0] = batch_stage1_measure_wave<P, 2>(f, out, a, s, n, cz, w, lds, lane); break;
                default: r[0] = batch_stage1_measure_wave<P, 3>(f, out, a, s, n, cz, w, lds, lane); break;
            }
        }
        __syncthreads();
        if (valid) clover_pair_terms(lds, lane, (int)threadIdx.y, r[1], r[2]);
        block_sum<3>(r, red, 4);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            partials[3 * tile + 0] = r[0];
            partials[3 * tile + 1] = r[1];
            partials[3 * tile + 2] = r[2];
        }
        __syncthreads();  // lds and red are reused by the next tile
    }
    if (w) atomicOr(warn, 1);
}

// One workgroup per configuration: ordered, compensated sum of its tiles.
__global__ __launch_bounds__(256) void k_batch_reduce(const double* __restrict__ partials,
                                                      const cfg_desc* __restrict__ cfgs, int ncfg,
                                                      double* __restrict__ out,
                                                      const unsigned* __restrict__ slot_ctr) {
    __shared__ double lds[4 * 3];
    const cfg_desc& c = cfgs[blockIdx.x];
    double v[3] = {0.0, 0.0, 0.0}, comp[3] = {0.0, 0.0, 0.0};
    strided_compensated_sum<3>(partials + 3 * (size_t)c.tile0, c.ntiles, threadIdx.x, blockDim.x, v,
                               comp);
#pragma unroll
    for (int k = 0; k < 3; k++) v[k] += comp[k];
    block_sum<3>(v, lds, 4);
    if (threadIdx.x == 0) {
        const unsigned slot = slot_ctr ? *slot_ctr : 0u;
        double* o = out + 3 * ((size_t)slot * ncfg + blockIdx.x);
        const double invV = 1.0 / (double)c.g.V;
        o[0] = v[0] / 3.0 / 6.0 * invV;
        o[1] = v[1] * invV;
        o[2] = v[2] * (1.0 / (4.0 * M_PI * M_PI));
    }
}

__global__ void k_bump(unsigned* ctr) { *ctr += 1u; }
__global__ void k_set(unsigned* ctr, unsigned v) { *ctr = v; }

hipError_t launch_set_counter(unsigned* ctr, unsigned value, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    k_set<<<1, 1, 0, st>>>(ctr, value);
    return hipGetLastError();
}

hipError_t launch_batch_measure(prec p, const void* U, const cfg_desc* cfgs,
                                const unsigned* tile_cfg, unsigned ntiles, int ncfg,
                                const batch_launch& bl, double* partials, double* out,
                                unsigned* slot_ctr, bool bump, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const dim3 block(kSitesPerBlock, 4);
    if (p == prec::f64)
        k_batch_measure<prec::f64><<<bl.grid, block, 0, st>>>(
            static_cast<const double2*>(U), cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win,
            partials);
    else
        k_batch_measure<prec::f32m><<<bl.grid, block, 0, st>>>(
            static_cast<const float4*>(U), cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win,
            partials);
    k_batch_reduce<<<ncfg, 256, 0, st>>>(partials, cfgs, ncfg, out, slot_ctr);
    if (bump && slot_ctr) k_bump<<<1, 1, 0, st>>>(slot_ctr);
    return hipGetLastError();
}

hipError_t launch_batch_stage1_measure(prec p, const void* Uin, void* Uout, void* acc,
                                       const cfg_desc* cfgs, const unsigned* tile_cfg,
                                       unsigned ntiles, int ncfg, const batch_launch& bl,
                                       double cz, int* warn, double* partials, double* out,
                                       unsigned* slot_ctr, bool bump, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const dim3 block(kSitesPerBlock, 4);
    if (p == prec::f64)
        k_batch_stage1_measure<prec::f64><<<bl.grid, block, 0, st>>>(
            static_cast<const double2*>(Uin), static_cast<double2*>(Uout),
            static_cast<double2*>(acc), cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, warn,
            partials);
    else
        k_batch_stage1_measure<prec::f32m><<<bl.grid, block, 0, st>>>(
            static_cast<const float4*>(Uin), static_cast<float4*>(Uout), static_cast<float4*>(acc),
            cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, warn, partials);
    k_batch_reduce<<<ncfg, 256, 0, st>>>(partials, cfgs, ncfg, out, slot_ctr);
    if (bump && slot_ctr) k_bump<<<1, 1, 0, st>>>(slot_ctr);
    return hipGetLastError();
}

}  // namespace qcb
