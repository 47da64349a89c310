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

}  // namespace qcb
