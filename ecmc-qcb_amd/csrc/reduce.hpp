// reduce.hpp -- deterministic wavefront / workgroup reductions.
//
// The reference sums observables with an OpenMP reduction plus MPI_Allreduce
// (src/observables/observables_mpi.cpp:115-156), whose order depends on the
// thread and rank counts.  Here the order is fixed by the launch shape: a
// butterfly inside each 64-wide wavefront, the wavefronts of a workgroup in
// index order, then the per-workgroup partials in index order, so repeated
// runs give bit-identical sums.
#pragma once
#include <hip/hip_runtime.h>

namespace qcb {

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Sum NV values over a workgroup of `nwaves` wavefronts; thread 0 of the
// workgroup returns the totals in v[].  lds must hold nwaves*NV doubles.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* lds, int nwaves) {
    const int lane = threadIdx.x & 63;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) lds[wave * NV + k] = v[k];
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = lds[k];
            for (int w = 1; w < nwaves; w++) s += lds[w * NV + k];
            v[k] = s;
        }
    }
}

// Neumaier's compensated summation step (Z. Angew. Math. Mech. 54 (1974) 39).
__device__ inline void neumaier_add(double& sum, double& comp, double x) {
    const double t = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
}

// Compensated sums over i = t, t + stride, t + 2 stride, ... < n of the NV
// values p[NV*i + k], in that order.  The loads of 8 consecutive terms are
// issued before their additions, so the loop waits on memory once per 8
// terms instead of once per term; the order of the additions is unchanged.
template <int NV>
__device__ inline void strided_compensated_sum(const double* __restrict__ p, unsigned n,
                                               unsigned t, unsigned stride, double (&v)[NV],
                                               double (&c)[NV]) {
    constexpr unsigned kB = 8;
    unsigned i = t;
    for (; i + (kB - 1) * stride < n; i += kB * stride) {
        double x[kB][NV];
#pragma unroll
        for (unsigned j = 0; j < kB; j++)
#pragma unroll
            for (int k = 0; k < NV; k++) x[j][k] = p[NV * (i + j * stride) + k];
#pragma unroll
        for (unsigned j = 0; j < kB; j++)
#pragma unroll
            for (int k = 0; k < NV; k++) neumaier_add(v[k], c[k], x[j][k]);
    }
    for (; i < n; i += stride)
#pragma unroll
        for (int k = 0; k < NV; k++) neumaier_add(v[k], c[k], p[NV * i + k]);
}

// Threads of the workgroup that reduces a lattice's per-tile partial triples
// (k_reduce_measure, k_batch_reduce).  Each thread adds every kReduceThreads-th
// triple: at 32^4 (16384 tiles) it took 11.8 us per measurement against
// 13.9 us with 256 threads.
constexpr int kReduceThreads = 1024;

// The ordered, compensated total of the n triples p[3i..3i+2] by one workgroup
// of kReduceThreads threads (thread t = linear thread index adds triples t,
// t + kReduceThreads, ...); thread 0 returns it in v[].  The single-lattice
// and ensemble reductions both use it, so an ensemble member's sums are
// bit-identical to a single lattice's.
__device__ inline void tile_triples_sum(const double* __restrict__ p, unsigned n, double (&v)[3]) {
    __shared__ double lds[kReduceThreads / 64 * 3];
    const unsigned tid = threadIdx.x + threadIdx.y * blockDim.x;
    double c[3] = {0.0, 0.0, 0.0};
    v[0] = v[1] = v[2] = 0.0;
    strided_compensated_sum<3>(p, n, tid, kReduceThreads, v, c);
#pragma unroll
    for (int k = 0; k < 3; k++) v[k] += c[k];
    block_sum<3>(v, lds, kReduceThreads / 64);
}

// The reference's normalisations of the summed triple: P = sum Re tr / 3 /
// (6V) (observables_mpi.cpp:48,67), E = sum e / V (:173-176),
// Q = sum q / (4 pi^2) (:132).
__device__ inline void write_observables(const double (&v)[3], double invV, double* out) {
    out[0] = v[0] / 3.0 / 6.0 * invV;
    out[1] = v[1] * invV;
    out[2] = v[2] * (1.0 / (4.0 * M_PI * M_PI));
}

}  // namespace qcb
