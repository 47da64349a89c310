// heatbath.hip -- Cabibbo-Marinari heatbath on the device: the step before the
// flow measurement (thermalised inputs for the flow, generated in HBM).
//
// Restates heatbath::generate_su2_kp (src/heatbath/heatbath.cpp:8-36),
// su2_step (:39-44) and mpi::heatbathcb::hit (src/heatbath/heatbath_mpi.cpp:
// 10-33): for each link, the staple A, then for the SU(2) subgroups (0,1),
// (1,2), (0,2): W = U A, w = the quaternion of W's 2x2 block, an SU(2) draw
// x at alpha = 2 beta |w| / 3 (su2_draw), U <- embed(x (w/|w|)^dag) U.
// The sweep order differs from the reference on purpose: one launch updates
// all U_mu(x) of one site parity for one mu, and such links share no staple
// on a lattice with even extents, so the update is race-free (the
// reference's OpenMP sweep over all mu of a parity is not).  Random numbers
// come from a counter-based generator keyed by (seed, sweep, mu, site,
// draw), so a run is reproducible on any device and launch shape.
#include "flow_body.hpp"
#include "kernels.hpp"
#include "su2_draw.hpp"

namespace qcb {

template <int MU>
__device__ inline void hb_link(double2* __restrict__ U, const lattice& g, unsigned s,
                               double beta, hb_rng& r, int& stuck) {
    const field64 f{U, g.V};
    const nbr n = neighbours(g, s);
    const mat3<double> A = staple_sum<field64, MU>(f, s, n);
    mat3<double> u = f.link(MU, s);
    const int sub[3][2] = {{0, 1}, {1, 2}, {0, 2}};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int i = sub[q][0], j = sub[q][1];
        const mat3<double> W = mul(u, A);
        const cpx<double> w00 = W.e[i][i], w01 = W.e[i][j], w10 = W.e[j][i], w11 = W.e[j][j];
        const double w[4] = {(w00.r + w11.r) / 2.0, (w01.i + w10.i) / 2.0, (w01.r - w10.r) / 2.0,
                             (w00.i - w11.i) / 2.0};
        const double k = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
        double x[4];
        if (!(k > 0.0) || !su2_draw(beta * k * 2.0 / 3.0, r, x)) {
            stuck = 1;
            continue;
        }
        const double v[4] = {w[0] / k, -w[1] / k, -w[2] / k, -w[3] / k};  // (w/|w|)^dag
        const double p[4] = {x[0] * v[0] - x[1] * v[1] - x[2] * v[2] - x[3] * v[3],
                             x[0] * v[1] + x[1] * v[0] + x[2] * v[3] - x[3] * v[2],
                             x[0] * v[2] - x[1] * v[3] + x[2] * v[0] + x[3] * v[1],
                             x[0] * v[3] + x[1] * v[2] - x[2] * v[1] + x[3] * v[0]};
        // U <- R U with R = embed(p) in rows/columns (i, j)
        const cpx<double> Rii = {p[0], p[3]}, Rjj = {p[0], -p[3]};
        const cpx<double> Rij = {p[2], p[1]}, Rji = {-p[2], p[1]};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const cpx<double> ui = u.e[i][c], uj = u.e[j][c];
            u.e[i][c] = madd(Rii * ui, Rij, uj);
            u.e[j][c] = madd(Rji * ui, Rjj, uj);
        }
    }
    store_link64(U, g.V, MU, s, u);
}

__global__ __launch_bounds__(256) void k_heatbath(double2* __restrict__ U, lattice g, int mu,
                                                  int parity, double beta, uint64_t seed,
                                                  unsigned sweep, int* __restrict__ stuck) {
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.V) return;
    unsigned r = s;
    const unsigned x = r % (unsigned)g.L[0];
    r /= (unsigned)g.L[0];
    const unsigned y = r % (unsigned)g.L[1];
    r /= (unsigned)g.L[1];
    const unsigned z = r % (unsigned)g.L[2];
    const unsigned t = r / (unsigned)g.L[2];
    if (((x + y + z + t) & 1u) != (unsigned)parity) return;
    hb_rng rng{hb_mix(hb_mix(seed ^ (0x632be59bd9b4e019ULL * (sweep + 1))) + 4ull * s + mu), 0u};
    int st = 0;
    switch (mu) {
        case 0: hb_link<0>(U, g, s, beta, rng, st); break;
        case 1: hb_link<1>(U, g, s, beta, rng, st); break;
        case 2: hb_link<2>(U, g, s, beta, rng, st); break;
        default: hb_link<3>(U, g, s, beta, rng, st); break;
    }
    if (st) atomicOr(stuck, 1);
}

// Column Gram-Schmidt re-unitarisation (GaugeField::projection_su3,
// src/gauge/GaugeField.cpp:36-58), applied after every sweep.
__global__ __launch_bounds__(256) void k_project_su3(double2* __restrict__ U, unsigned V) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= V) return;
    const int mu = threadIdx.y;
    mat3<double> m = field64{U, V}.link(mu, s);
    cpx<double> c0[3], c1[3];
    for (int i = 0; i < 3; i++) {
        c0[i] = m.e[i][0];
        c1[i] = m.e[i][1];
    }
    double n0 = norm2(c0[0]) + norm2(c0[1]) + norm2(c0[2]);
    const double i0 = 1.0 / sqrt(n0);
    for (int i = 0; i < 3; i++) c0[i] = i0 * c0[i];
    cpx<double> dot = {0.0, 0.0};
    for (int i = 0; i < 3; i++) dot = madd_cl(dot, c0[i], c1[i]);
    for (int i = 0; i < 3; i++) c1[i] = c1[i] - dot * c0[i];
    const double n1 = norm2(c1[0]) + norm2(c1[1]) + norm2(c1[2]);
    const double i1 = 1.0 / sqrt(n1);
    for (int i = 0; i < 3; i++) c1[i] = i1 * c1[i];
    for (int i = 0; i < 3; i++) {
        m.e[i][0] = c0[i];
        m.e[i][1] = c1[i];
    }
    m.e[0][2] = conj(c0[1] * c1[2] - c0[2] * c1[1]);
    m.e[1][2] = conj(c0[2] * c1[0] - c0[0] * c1[2]);
    m.e[2][2] = conj(c0[0] * c1[1] - c0[1] * c1[0]);
    store_link64(U, V, mu, s, m);
}

hipError_t launch_heatbath_f64(double2* U, const lattice& g, double beta, int n_sweeps,
                               uint64_t seed, unsigned sweep0, int* stuck, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const unsigned nb = (g.V + 255) / 256;
    for (int sw = 0; sw < n_sweeps; sw++) {
        for (int mu = 0; mu < 4; mu++)
            for (int par = 0; par < 2; par++)
                k_heatbath<<<nb, 256, 0, st>>>(U, g, mu, par, beta, seed, sweep0 + sw, stuck);
        k_project_su3<<<(g.V + kSitesPerBlock - 1) / kSitesPerBlock, dim3(kSitesPerBlock, 4), 0,
                        st>>>(U, g.V);
    }
    return hipGetLastError();
}

}  // namespace qcb
