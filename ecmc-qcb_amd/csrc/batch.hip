// batch.hip -- ensemble kernels: one launch per RK3 stage (and one
// measurement + one reduction per measurement) for a whole ensemble of
// lattices of uneven volumes.
//
// The persistent grid is sized to the resident capacity of the GPU; workgroup b
// visits virtual slots v = b, b + grid, b + 2 grid, ... and each slot maps to a
// tile of the flat (configuration, tile) list through the XCD-aware order of
// tile_of_block.  With a grid of zero the same kernels run one workgroup per
// slot instead (for_each_slot).  The per-link and per-site work is exactly that of the
// single-lattice kernels (flow_body.hpp, measure_body.hpp).
#include "batch.hpp"
#include "flow_body.hpp"
#include "kernels.hpp"
#include "measure_body.hpp"
#include "reduce.hpp"
#include "stage_lds.hpp"

namespace qcb {

// Slots of this workgroup: b, b + grid, b + 2 grid, ... on the persistent
// grid, or the one slot b when the grid has a workgroup per slot.  The flat
// form exists because the loop of the persistent form costs the register
// allocator up to 40 VGPRs in these kernels (a wavefront per SIMD); which of
// the two is faster for an ensemble is the autotuner's call.  The per-slot
// bodies are functions with __restrict__ pointer parameters, so the loads of
// the input field may move past the stores of the output as in the
// single-lattice kernels.
#define QCB_FOR_EACH_SLOT(PERSIST, nslots, call)                             \
    do {                                                                     \
        if constexpr (PERSIST) {                                             \
            for (unsigned v = blockIdx.x; v < (nslots); v += gridDim.x) call; \
        } else if (blockIdx.x < (nslots)) {                                  \
            const unsigned v = blockIdx.x;                                   \
            call;                                                            \
        }                                                                    \
    } while (0)

// A wavefront's 64 sites are one tile of one configuration, so that
// configuration's field offsets and volume are the same for all its lanes:
// taken into scalar registers, the field's word bases stay wave-uniform and
// the loads take the site as a 32-bit offset (fields.hpp: word_at) instead
// of a 64-bit address per lane and word.
__device__ inline size_t wave_uniform(size_t v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((size_t)hi << 32) | lo;
}
__device__ inline unsigned wave_uniform(unsigned v) {
    return __builtin_amdgcn_readfirstlane(v);
}

template <prec P, int STAGE, int DPW>
__device__ __forceinline__ void batch_stage_slot(unsigned v,
                                                 const typename storage<P>::vec* __restrict__ Uin,
                                                 typename storage<P>::vec* __restrict__ Uout,
                                                 typename storage<P>::avec* __restrict__ acc,
                                                 const cfg_desc* __restrict__ cfgs,
                                                 const unsigned* __restrict__ tile_cfg,
                                                 unsigned ntiles, int map_mode, unsigned map_win,
                                                 double cz, double ca, bool& w) {
    // slot v is a group of DPW tiles (group_tiles); 4/DPW wavefronts per tile
    const tile_groups G = group_tiles(ntiles, 0u, map_mode, map_win, DPW);
    const unsigned tile = group_tile<DPW>(tile_of_block(v, G.n, map_mode, G.win), threadIdx.y);
    if (tile >= ntiles) return;
    const cfg_desc& c = cfgs[tile_cfg[tile]];
    unsigned co[4];
    const unsigned s = tile_site(c.g, tile - c.tile0, threadIdx.x, co);
    if (s >= c.g.V) return;
    const nbr n = neighbours_at(c.g, co);
    const size_t uo = wave_uniform(c.u_off), ao = wave_uniform(c.a_off);
    stage_directions<P, STAGE, false, DPW>(threadIdx.y, Uin + uo, Uout + uo, acc + ao,
                                           wave_uniform(c.g.V), s, n, cz, ca, w);
}

// Wavefronts per SIMD of the one-workgroup-per-tile stage kernel, which the
// register allocator reaches here only when asked (FP32 with 2 or 4
// directions per wavefront: 73-77 VGPRs instead of 120-129).  FP32 with one
// direction per wavefront takes 5: at 6 (80 VGPRs) it spilled 4 dwords to
// scratch memory and ran its ensemble 2% slower than at 5 (86 VGPRs).
template <prec P, int DPW>
constexpr int kFlatStageWaves = P == prec::f64 ? 3 : DPW == 1 ? 5 : 6;

template <prec P, int STAGE, bool PERSIST, int DPW>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(PERSIST ? 1 : kFlatStageWaves<P, DPW>))) void k_batch_stage(const typename storage<P>::vec* __restrict__ Uin,
                                                     typename storage<P>::vec* __restrict__ Uout,
                                                     typename storage<P>::avec* __restrict__ acc,
                                                     const cfg_desc* __restrict__ cfgs,
                                                     const unsigned* __restrict__ tile_cfg,
                                                     unsigned ntiles, int map_mode, unsigned map_win,
                                                     double cz, double ca, int* __restrict__ warn) {
    bool w = false;
    const tile_groups G = group_tiles(ntiles, 0u, map_mode, map_win, DPW);
    const unsigned nslots = grid_for_tiles(G.n, map_mode, G.win);
    QCB_FOR_EACH_SLOT(PERSIST, nslots,
                      (batch_stage_slot<P, STAGE, DPW>(v, Uin, Uout, acc, cfgs, tile_cfg, ntiles,
                                                       map_mode, map_win, cz, ca, w)));
    if (w) atomicOr(warn, 1);
}

// One workgroup per 64-site tile with the links its four direction
// wavefronts share staged in LDS (k_rk3_stage_lds without whole x-rows, so
// the same body for every member): needs every member's volume a multiple of
// 64 with 64-consecutive-site tiles, one thread per link and the flat grid.
// The products are those of the plain body in the same order, so the flow
// stays bit-identical to it and to the single-lattice path.
template <prec P, int STAGE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kLdsStageWaves<P>))) void
k_batch_stage_lds(const typename storage<P>::vec* __restrict__ Uin,
                  typename storage<P>::vec* __restrict__ Uout,
                  typename storage<P>::avec* __restrict__ acc, const cfg_desc* __restrict__ cfgs,
                  const unsigned* __restrict__ tile_cfg, unsigned ntiles, int map_mode,
                  unsigned map_win, double cz, double ca, int* __restrict__ warn) {
    __shared__ lds_link<P> own[4], dn[4];
    const unsigned tile = tile_of_block(blockIdx.x, ntiles, map_mode, map_win);
    if (tile >= ntiles) return;  // the whole workgroup: no barrier is skipped
    const cfg_desc& c = cfgs[tile_cfg[tile]];
    const unsigned lane = threadIdx.x;
    const int mu = (int)threadIdx.y;  // wavefront-uniform
    unsigned co[4];
    const unsigned s = tile_site(c.g, tile - c.tile0, lane, co);  // full tiles: s < V
    const nbr n = neighbours_at(c.g, co);
    const size_t uo = wave_uniform(c.u_off), ao = wave_uniform(c.a_off);
    const unsigned V = wave_uniform(c.g.V);
    stage_tile_to_lds<P, false>(Uin + uo, V, mu, s, n, own, dn);
    __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0): the LDS copies have landed
    __syncthreads();
    const auto f = storage<P>::view(Uin + uo, V);
    auto* out = Uout + uo;
    auto* a = acc + ao;
    bool w = false;
    switch (mu) {
        case 0: stage_lds_body<P, STAGE, 0, false>(f, out, a, s, n, own, dn, lane, cz, ca, w); break;
        case 1: stage_lds_body<P, STAGE, 1, false>(f, out, a, s, n, own, dn, lane, cz, ca, w); break;
        case 2: stage_lds_body<P, STAGE, 2, false>(f, out, a, s, n, own, dn, lane, cz, ca, w); break;
        default: stage_lds_body<P, STAGE, 3, false>(f, out, a, s, n, own, dn, lane, cz, ca, w); break;
    }
    if (w) atomicOr(warn, 1);
}

template <prec P, int STAGE, int DPW>
static void batch_stage_kernel(const batch_launch& bl, hipStream_t st,
                               const typename storage<P>::vec* in, typename storage<P>::vec* out,
                               typename storage<P>::avec* a, const cfg_desc* cfgs,
                               const unsigned* tile_cfg, unsigned ntiles, double cz, double ca,
                               int* warn) {
    const dim3 block(kSitesPerBlock, 4);
    const tile_groups G = group_tiles(ntiles, 0u, bl.map_mode, bl.map_win, DPW);
    const unsigned nslots = grid_for_tiles(G.n, bl.map_mode, G.win);
    if (bl.grid)
        k_batch_stage<P, STAGE, true, DPW><<<bl.grid, block, 0, st>>>(
            in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, ca, warn);
    else
        k_batch_stage<P, STAGE, false, DPW><<<nslots, block, 0, st>>>(
            in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, ca, warn);
}

template <prec P, int STAGE>
static hipError_t batch_stage_kernel(const batch_launch& bl, hipStream_t st,
                                     const typename storage<P>::vec* in,
                                     typename storage<P>::vec* out, typename storage<P>::avec* a,
                                     const cfg_desc* cfgs, const unsigned* tile_cfg,
                                     unsigned ntiles, double cz, double ca, int* warn) {
    if (bl.lds) {  // flat grid, one thread per link (the host checks)
        k_batch_stage_lds<P, STAGE>
            <<<grid_for_tiles(ntiles, bl.map_mode, bl.map_win), dim3(kSitesPerBlock, 4), 0, st>>>(
                in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, ca, warn);
        return hipSuccess;
    }
    switch (bl.dpw) {
        case 1: batch_stage_kernel<P, STAGE, 1>(bl, st, in, out, a, cfgs, tile_cfg, ntiles, cz, ca, warn); break;
        case 2: batch_stage_kernel<P, STAGE, 2>(bl, st, in, out, a, cfgs, tile_cfg, ntiles, cz, ca, warn); break;
        case 4: batch_stage_kernel<P, STAGE, 4>(bl, st, in, out, a, cfgs, tile_cfg, ntiles, cz, ca, warn); break;
        default: return hipErrorInvalidValue;
    }
    return hipSuccess;
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
    hipError_t e;
    switch (stage) {
        case 1: e = batch_stage_kernel<P, 1>(bl, st, in, out, a, cfgs, tile_cfg, ntiles, cz, ca, warn); break;
        case 2: e = batch_stage_kernel<P, 2>(bl, st, in, out, a, cfgs, tile_cfg, ntiles, cz, ca, warn); break;
        case 3: e = batch_stage_kernel<P, 3>(bl, st, in, out, a, cfgs, tile_cfg, ntiles, cz, ca, warn); break;
        default: return hipErrorInvalidValue;
    }
    return e != hipSuccess ? e : hipGetLastError();
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
__device__ __forceinline__ void batch_measure_slot(unsigned v,
                                                   const typename storage<P>::vec* __restrict__ U,
                                                   const cfg_desc* __restrict__ cfgs,
                                                   const unsigned* __restrict__ tile_cfg,
                                                   unsigned ntiles, int map_mode, unsigned map_win,
                                                   double* __restrict__ partials,
                                                   typename storage<P>::real* lds, double* red) {
    using FT = typename storage<P>::field;
    const unsigned tile = tile_of_block(v, ntiles, map_mode, map_win);
    if (tile >= ntiles) return;  // uniform per workgroup: no barrier is skipped by part of it
    const unsigned lane = threadIdx.x;
    const cfg_desc& c = cfgs[tile_cfg[tile]];
    const unsigned s = (tile - c.tile0) * kSitesPerBlock + lane;
    const bool valid = s < c.g.V;
    const FT f = storage<P>::view(U + wave_uniform(c.u_off), wave_uniform(c.g.V));
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

template <prec P, bool PERSIST>
__global__ __launch_bounds__(256, 2) void k_batch_measure(const typename storage<P>::vec* __restrict__ U,
                                                          const cfg_desc* __restrict__ cfgs,
                                                          const unsigned* __restrict__ tile_cfg,
                                                          unsigned ntiles, int map_mode,
                                                          unsigned map_win,
                                                          double* __restrict__ partials) {
    __shared__ typename storage<P>::real lds[kCloverSlots * 9 * 64];
    __shared__ double red[4 * 3];
    const unsigned nslots = grid_for_tiles(ntiles, map_mode, map_win);
    QCB_FOR_EACH_SLOT(PERSIST, nslots,
                      (batch_measure_slot<P>(v, U, cfgs, tile_cfg, ntiles, map_mode, map_win,
                                             partials, lds, red)));
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
__device__ __forceinline__ void batch_stage1_measure_slot(
    unsigned v, const typename storage<P>::vec* __restrict__ Uin,
    typename storage<P>::vec* __restrict__ Uout, typename storage<P>::avec* __restrict__ acc,
    const cfg_desc* __restrict__ cfgs, const unsigned* __restrict__ tile_cfg, unsigned ntiles,
    int map_mode, unsigned map_win, double cz, bool& w, double* __restrict__ partials,
    typename storage<P>::real* lds, double* red) {
    using FT = typename storage<P>::field;
    const unsigned tile = tile_of_block(v, ntiles, map_mode, map_win);
    if (tile >= ntiles) return;  // uniform per workgroup: no barrier is skipped by part of it
    const unsigned lane = threadIdx.x;
    const cfg_desc& c = cfgs[tile_cfg[tile]];
    const unsigned s = (tile - c.tile0) * kSitesPerBlock + lane;
    const bool valid = s < c.g.V;
    const size_t uo = wave_uniform(c.u_off);
    const FT f = storage<P>::view(Uin + uo, wave_uniform(c.g.V));
    const nbr n = neighbours(c.g, valid ? s : 0u);
    auto* out = Uout + uo;
    auto* a = acc + wave_uniform(c.a_off);
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

template <prec P, bool PERSIST>
__global__ __launch_bounds__(256, 2) void k_batch_stage1_measure(
    const typename storage<P>::vec* __restrict__ Uin, typename storage<P>::vec* __restrict__ Uout,
    typename storage<P>::avec* __restrict__ acc, const cfg_desc* __restrict__ cfgs,
    const unsigned* __restrict__ tile_cfg, unsigned ntiles, int map_mode, unsigned map_win,
    double cz, int* __restrict__ warn, double* __restrict__ partials) {
    if constexpr (!kFusedMeasureBuilt) return;  // never launched (host checks)
    __shared__ typename storage<P>::real lds[kCloverSlots * 9 * 64];
    __shared__ double red[4 * 3];
    bool w = false;
    const unsigned nslots = grid_for_tiles(ntiles, map_mode, map_win);
    QCB_FOR_EACH_SLOT(PERSIST, nslots,
                      (batch_stage1_measure_slot<P>(v, Uin, Uout, acc, cfgs, tile_cfg, ntiles,
                                                    map_mode, map_win, cz, w, partials, lds,
                                                    red)));
    if (w) atomicOr(warn, 1);
}

// One workgroup per configuration: ordered, compensated sum of its tiles.
__global__ __launch_bounds__(kReduceThreads) void k_batch_reduce(
    const double* __restrict__ partials, const cfg_desc* __restrict__ cfgs, int ncfg,
    double* __restrict__ out, const unsigned* __restrict__ slot_ctr) {
    const cfg_desc& c = cfgs[blockIdx.x];
    double v[3];
    tile_triples_sum(partials + 3 * (size_t)c.tile0, c.ntiles, v);
    if (threadIdx.x == 0) {
        const unsigned slot = slot_ctr ? *slot_ctr : 0u;
        write_observables(v, 1.0 / (double)c.g.V, out + 3 * ((size_t)slot * ncfg + blockIdx.x));
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
    const unsigned nslots = grid_for_tiles(ntiles, bl.map_mode, bl.map_win);
    const auto* u64 = static_cast<const double2*>(U);
    const auto* u32 = static_cast<const float4*>(U);
    if (p == prec::f64 && bl.grid)
        k_batch_measure<prec::f64, true><<<bl.grid, block, 0, st>>>(
            u64, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, partials);
    else if (p == prec::f64)
        k_batch_measure<prec::f64, false><<<nslots, block, 0, st>>>(
            u64, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, partials);
    else if (bl.grid)
        k_batch_measure<prec::f32m, true><<<bl.grid, block, 0, st>>>(
            u32, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, partials);
    else
        k_batch_measure<prec::f32m, false><<<nslots, block, 0, st>>>(
            u32, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, partials);
    k_batch_reduce<<<ncfg, kReduceThreads, 0, st>>>(partials, cfgs, ncfg, out, slot_ctr);
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
    const unsigned nslots = grid_for_tiles(ntiles, bl.map_mode, bl.map_win);
    const unsigned grid = bl.grid ? bl.grid : nslots;
    if (p == prec::f64) {
        const auto* in = static_cast<const double2*>(Uin);
        auto* out = static_cast<double2*>(Uout);
        auto* a = static_cast<double2*>(acc);
        if (bl.grid)
            k_batch_stage1_measure<prec::f64, true><<<grid, block, 0, st>>>(
                in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, warn, partials);
        else
            k_batch_stage1_measure<prec::f64, false><<<grid, block, 0, st>>>(
                in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, warn, partials);
    } else {
        const auto* in = static_cast<const float4*>(Uin);
        auto* out = static_cast<float4*>(Uout);
        auto* a = static_cast<float4*>(acc);
        if (bl.grid)
            k_batch_stage1_measure<prec::f32m, true><<<grid, block, 0, st>>>(
                in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, warn, partials);
        else
            k_batch_stage1_measure<prec::f32m, false><<<grid, block, 0, st>>>(
                in, out, a, cfgs, tile_cfg, ntiles, bl.map_mode, bl.map_win, cz, warn, partials);
    }
    k_batch_reduce<<<ncfg, kReduceThreads, 0, st>>>(partials, cfgs, ncfg, out, slot_ctr);
    if (bump && slot_ctr) k_bump<<<1, 1, 0, st>>>(slot_ctr);
    return hipGetLastError();
}

}  // namespace qcb
