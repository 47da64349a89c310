// kernels.hpp -- host-side launchers of the HIP kernels (internal to libqcbflow).
//
// Link and su(3) field pointers are untyped here; their layout is fixed by the
// precision `p` (fields.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fields.hpp"

namespace qcb {

// Sites per workgroup along the lexicographic order; each workgroup runs the
// four directions mu of those sites as four wavefronts.
constexpr int kSitesPerBlock = 64;

// Number of per-workgroup partial triples written by the measurement kernel.
__host__ __device__ inline unsigned measure_blocks(const lattice& g) {
    return (g.V + kSitesPerBlock - 1) / kSitesPerBlock;
}

// ---- flow ------------------------------------------------------------------------
// One fused RK3 stage (1, 2 or 3): staples -> -P_TA(U Sigma) -> RK3
// combination with the su(3) accumulator -> exp -> exp(X) U into Uout.
//   stage 1: X = cz Z,          acc <- Z
//   stage 2: X = cz Z - ca acc, acc <- 8/9 Z - 17/36 acc
//   stage 3: X = cz Z - ca acc
hipError_t launch_rk3_stage(prec p, int stage, const void* Uin, void* Uout, void* acc,
                            const lattice& g, double cz, double ca, int* warn, hipStream_t st);

// The stage kernel with the links shared by the four direction wavefronts of
// a tile staged in LDS (stage_lds.hip): the same flow, bit for bit.  Applies
// to 64-consecutive-site tiles, one thread per link and V a multiple of 64
// (FP32-mixed: also whole x-rows per wavefront).
bool stage_lds_applies(prec p, const lattice& g);
hipError_t launch_rk3_stage_lds(prec p, int stage, const void* Uin, void* Uout, void* acc,
                                const lattice& g, double cz, double ca, int* warn,
                                hipStream_t st);

// The fused measurement hooks into the SU(3)-product staple sum: built unless
// the library is compiled with QCB_SU3_PRODUCTS=0 or with streamed FP64 rows.
constexpr bool kFusedMeasureBuilt = QCB_SU3_PRODUCTS && !field64::stream_rows;

// RK3 stage 1 fused with the clover measurement of Uin: the stage as above,
// plus the measurement's per-tile partials and their reduction into
// out[0..2] = {P, E, Q}, bit-identical to launch_measure(clover = true).
// Needs the default stage tiles (64 consecutive sites).
hipError_t launch_rk3_stage1_measure(prec p, const void* Uin, void* Uout, void* acc,
                                     const lattice& g, double cz, int* warn, double* partials,
                                     double* out, hipStream_t st);

// Unfused pieces, one per GradientFlow member function.
hipError_t launch_force(prec p, const void* U, void* F, const lattice& g, hipStream_t st);
// which = 1: U <- exp(eps/4 F0) U ; 2: exp(8/9 eps F1 - 17/36 eps F0) U ;
// 3: exp(3/4 eps F1 - eps F0) U.  In place.
hipError_t launch_update(prec p, int which, void* U, const void* F0, const void* F1,
                         const lattice& g, double eps, int* warn, hipStream_t st);
hipError_t launch_replace_force0(prec p, void* F0, const void* F1, const lattice& g,
                                 hipStream_t st);

// ---- observables -----------------------------------------------------------------
// Per-workgroup partial sums (plaquette, E, Q-density) written to
// partials[3*measure_blocks], then reduced in a fixed order into
// out[0..2] = {P, E, Q} with the reference normalisations.  clover = false
// computes the plaquette only (out[1], out[2] = 0).
hipError_t launch_measure(prec p, const void* U, const lattice& g, double* partials, double* out,
                          bool clover, hipStream_t st);
// The ordered reduction of launch_measure alone (partials of measure_blocks(g)
// tiles -> out[0..2]).
hipError_t launch_reduce_measure(const double* partials, const lattice& g, double* out,
                                 hipStream_t st);
// clover_site (observables_mpi.cpp:66-90) at one site: out[0..17] = F_munu(x)
// = (C - C^dag)/8, row-major (re, im); zero for mu == nu.
hipError_t launch_clover_site(prec p, const void* U, const lattice& g, unsigned site, int mu,
                              int nu, double* out, hipStream_t st);
// local_q_e_clover (observables_mpi.cpp:93-112) at one site: out = {q(x), e(x)}.
hipError_t launch_local_q_e(prec p, const void* U, const lattice& g, unsigned site, double* out,
                            hipStream_t st);
// max over links of |U U^dag - 1|_max and of |det U - 1|, into out[0..1]
// (partials sized 2*measure_blocks)
hipError_t launch_unitarity(prec p, const void* U, const lattice& g, double* partials,
                            double* out, hipStream_t st);

// ---- layout conversion and initialisation ----------------------------------------
// nsites sites starting at site0, 72 words per site in the reference / ILDG
// order (site, mu, i, j, re/im) <-> the device field.
//   kind: 0 native doubles, 1 big-endian doubles, 2 big-endian floats.
hipError_t launch_aos_to_field(prec p, const void* aos, int kind, void* U, const lattice& g,
                               unsigned site0, unsigned nsites, hipStream_t st);
hipError_t launch_field_to_aos(prec p, const void* U, double* aos, const lattice& g,
                               unsigned site0, unsigned nsites, bool bswap, hipStream_t st);
// nsites su(3) elements per direction from site0 on, as full 3x3 matrices in
// the link order (72 doubles per site).
hipError_t launch_alg_to_aos(prec p, const void* alg, double* aos, const lattice& g,
                             unsigned site0, unsigned nsites, hipStream_t st);
hipError_t launch_hot_start(prec p, void* U, const lattice& g, uint64_t seed, hipStream_t st);
hipError_t launch_cold_start(prec p, void* U, const lattice& g, hipStream_t st);
// out_mu(x) = in_mu(x - shift) on the periodic lattice, 0 <= shift[d] < L[d].
hipError_t launch_shift(prec p, const void* in, void* out, const lattice& g, const int shift[4],
                        hipStream_t st);
// Rebuild the third row of every link of an FP64 flow output in place (no-op
// for the FP32-mixed form, which never stores it).
hipError_t launch_complete_rows(prec p, void* U, unsigned V, hipStream_t st);

// n_sweeps checkerboard Cabibbo-Marinari heatbath sweeps at beta on an FP64
// field with even extents; sweep numbers sweep0.. key the random streams.
// *stuck is set if an SU(2) draw hit its retry bound.
hipError_t launch_heatbath_f64(double2* U, const lattice& g, double beta, int n_sweeps,
                               uint64_t seed, unsigned sweep0, int* stuck, hipStream_t st);

// exp test entry: n matrices Z (18 doubles each, AoS) -> exp(coeff * P(Z)) where
// P keeps the traceless anti-hermitian entries Z is assumed to have.
hipError_t launch_exp_batch_f64(const double* Z, double* out, int n, double coeff, int* warn,
                                hipStream_t st);

// Plain device copy (the bandwidth calibration kernel for traffic counters).
hipError_t launch_copy(const double2* src, double2* dst, size_t n, hipStream_t st);

}  // namespace qcb
