// layout.hip -- layout conversion, initialisation and utility kernels.
//
// Host buffers and ILDG payloads use the reference's interior order (site
// x-fastest, mu, 3x3 row-major, re/im: src/io/ildg.cpp:233-246); in HBM the
// links are structure-of-arrays (lattice.hpp).  The conversion stages a tile
// of 64 sites through LDS so both the AoS side and the SoA side are accessed
// with coalesced loads and stores; the big-endian byte swap of ILDG files
// (swap_double_endian, src/io/ildg.cpp:25-32) is fused into the same pass.
// For the FP32-mixed field the same pass keeps the first two rows, re-unitarises
// them in FP64 and rounds them to FP32; the way back rebuilds the third row in
// FP64.
#include "fields.hpp"
#include "kernels.hpp"

namespace qcb {

constexpr int kTile = 64;  // sites per conversion tile

__device__ inline double bswap_d(double v) {
    unsigned long long b = __double_as_longlong(v);
    return __longlong_as_double((long long)__builtin_bswap64(b));
}
__device__ inline float bswap_f(float v) {
    return __uint_as_float(__builtin_bswap32(__float_as_uint(v)));
}

// AoS (72 words per site) -> device field.  One workgroup per tile of 64 sites.
// kind: 0 native doubles, 1 big-endian doubles, 2 big-endian floats.
template <prec P>
__global__ __launch_bounds__(256) void k_aos_to_field(const void* __restrict__ aos,
                                                      typename storage<P>::vec* __restrict__ U,
                                                      unsigned V, unsigned site0, unsigned nsites,
                                                      int kind) {
    __shared__ double tile[kTile * 73];  // +1 pad per site row against bank conflicts
    const unsigned t0 = blockIdx.x * kTile;
    const unsigned nt = min((unsigned)kTile, nsites - t0);
    for (unsigned k = threadIdx.x; k < nt * 72; k += blockDim.x) {
        const size_t idx = (size_t)t0 * 72 + k;
        double d;
        if (kind == 2)
            d = (double)bswap_f(static_cast<const float*>(aos)[idx]);
        else {
            d = static_cast<const double*>(aos)[idx];
            if (kind == 1) d = bswap_d(d);
        }
        tile[(k / 72) * 73 + (k % 72)] = d;
    }
    __syncthreads();
    if constexpr (P == prec::f64) {
        // 36 complex entries (mu*9 + e) x nt sites; consecutive threads -> consecutive sites
        for (unsigned k = threadIdx.x; k < 36u * kTile; k += blockDim.x) {
            const unsigned site = k % kTile, comp = k / kTile;
            if (site < nt) {
                const double re = tile[site * 73 + 2 * comp], im = tile[site * 73 + 2 * comp + 1];
                U[(size_t)comp * V + site0 + t0 + site] = make_double2(re, im);
            }
        }
    } else {
        // one thread per link: rows 0 and 1, re-unitarised in FP64, to FP32
        const unsigned site = threadIdx.x % kTile, mu = threadIdx.x / kTile;
        if (site < nt) {
            const double* l = tile + site * 73 + mu * 18;
            cpx<double> r0[3], r1[3];
            for (int k = 0; k < 3; k++) {
                r0[k] = {l[2 * k], l[2 * k + 1]};
                r1[k] = {l[6 + 2 * k], l[6 + 2 * k + 1]};
            }
            store_rows32c(U, V, mu, site0 + t0 + site, r0, r1);
        }
    }
}

template <prec P>
__global__ __launch_bounds__(256) void k_field_to_aos(const typename storage<P>::vec* __restrict__ U,
                                                      double* __restrict__ aos, unsigned V,
                                                      unsigned site0, unsigned nsites, int bswap) {
    __shared__ double tile[kTile * 73];
    const unsigned t0 = blockIdx.x * kTile;
    const unsigned nt = min((unsigned)kTile, nsites - t0);
    if constexpr (P == prec::f64) {
        for (unsigned k = threadIdx.x; k < 36u * kTile; k += blockDim.x) {
            const unsigned site = k % kTile, comp = k / kTile;
            if (site < nt) {
                const double2 v = U[(size_t)comp * V + site0 + t0 + site];
                tile[site * 73 + 2 * comp] = v.x;
                tile[site * 73 + 2 * comp + 1] = v.y;
            }
        }
    } else {
        const unsigned site = threadIdx.x % kTile, mu = threadIdx.x / kTile;
        if (site < nt) {
            const float4* b = U + (size_t)mu * 3 * V + site0 + t0 + site;
            const float4 a = b[0], c = b[V], d = b[2 * (size_t)V];
            cpx<double> m[3][3];
            m[0][0] = {a.x, a.y};
            m[0][1] = {a.z, a.w};
            m[0][2] = {c.x, c.y};
            m[1][0] = {c.z, c.w};
            m[1][1] = {d.x, d.y};
            m[1][2] = {d.z, d.w};
            m[2][0] = conj(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
            m[2][1] = conj(m[0][2] * m[1][0] - m[0][0] * m[1][2]);
            m[2][2] = conj(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
            double* l = tile + site * 73 + mu * 18;
            for (int k = 0; k < 9; k++) {
                l[2 * k] = m[k / 3][k % 3].r;
                l[2 * k + 1] = m[k / 3][k % 3].i;
            }
        }
    }
    __syncthreads();
    double* dst = aos + (size_t)t0 * 72;
    for (unsigned k = threadIdx.x; k < nt * 72; k += blockDim.x) {
        const double d = tile[(k / 72) * 73 + (k % 72)];
        dst[k] = bswap ? bswap_d(d) : d;
    }
}

hipError_t launch_aos_to_field(prec p, const void* aos, int kind, void* U, const lattice& g,
                               unsigned site0, unsigned nsites, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (nsites == 0) return hipSuccess;
    const unsigned nb = (nsites + kTile - 1) / kTile;
    if (p == prec::f64)
        k_aos_to_field<prec::f64><<<nb, 256, 0, st>>>(aos, static_cast<double2*>(U), g.V, site0,
                                                      nsites, kind);
    else
        k_aos_to_field<prec::f32m><<<nb, 256, 0, st>>>(aos, static_cast<float4*>(U), g.V, site0,
                                                       nsites, kind);
    return hipGetLastError();
}

hipError_t launch_field_to_aos(prec p, const void* U, double* aos, const lattice& g,
                               unsigned site0, unsigned nsites, bool bswap, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (nsites == 0) return hipSuccess;
    const unsigned nb = (nsites + kTile - 1) / kTile;
    if (p == prec::f64)
        k_field_to_aos<prec::f64><<<nb, 256, 0, st>>>(static_cast<const double2*>(U), aos, g.V,
                                                      site0, nsites, bswap ? 1 : 0);
    else
        k_field_to_aos<prec::f32m><<<nb, 256, 0, st>>>(static_cast<const float4*>(U), aos, g.V,
                                                       site0, nsites, bswap ? 1 : 0);
    return hipGetLastError();
}

// su(3) field (8 reals per link) -> full 3x3 AoS doubles
template <prec P>
__global__ __launch_bounds__(256) void k_alg_to_aos(const typename storage<P>::avec* __restrict__ alg,
                                                    double* __restrict__ aos, unsigned V,
                                                    unsigned site0, unsigned nsites) {
    const unsigned i = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (i >= nsites) return;
    const unsigned s = site0 + i;
    const int mu = threadIdx.y;
    const auto m = alg_to_mat(storage<P>::load_alg(alg, V, mu, s));
    double* d = aos + ((size_t)i * 4 + mu) * 18;
    for (int k = 0; k < 9; k++) {
        d[2 * k] = (double)m.e[k / 3][k % 3].r;
        d[2 * k + 1] = (double)m.e[k / 3][k % 3].i;
    }
}
hipError_t launch_alg_to_aos(prec p, const void* alg, double* aos, const lattice& g,
                             unsigned site0, unsigned nsites, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (nsites == 0) return hipSuccess;
    if (site0 > g.V || nsites > g.V - site0) return hipErrorInvalidValue;
    const unsigned nb = (nsites + kSitesPerBlock - 1) / kSitesPerBlock;
    if (p == prec::f64)
        k_alg_to_aos<prec::f64><<<nb, dim3(kSitesPerBlock, 4), 0, st>>>(
            static_cast<const double2*>(alg), aos, g.V, site0, nsites);
    else
        k_alg_to_aos<prec::f32m><<<nb, dim3(kSitesPerBlock, 4), 0, st>>>(
            static_cast<const float4*>(alg), aos, g.V, site0, nsites);
    return hipGetLastError();
}

// ---- initialisation ---------------------------------------------------------
// Counter-based generator: splitmix64 finaliser of (seed, link, draw).
__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ inline double unif01(uint64_t seed, uint64_t link, uint32_t k) {
    const uint64_t h = mix64(mix64(seed + 0x9e3779b97f4a7c15ULL * (link + 1)) + k);
    return ((h >> 11) + 1) * 0x1.0p-53;  // (0, 1]
}

// Haar-random SU(3) per link: complex Gaussian matrix, column Gram-Schmidt,
// then division by det^{1/3} (the reference's hot_start uses a QR of a
// Gaussian matrix the same way: src/su3/utils.cpp:13-30).
template <prec P>
__global__ __launch_bounds__(256) void k_hot_start(typename storage<P>::vec* __restrict__ U,
                                                   unsigned V, uint64_t seed) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= V) return;
    const int mu = threadIdx.y;
    const uint64_t link = (uint64_t)s * 4 + mu;
    cpx<double> z[3][3];
    for (int k = 0; k < 9; k++) {
        const double u1 = unif01(seed, link, 2 * k), u2 = unif01(seed, link, 2 * k + 1);
        const double rad = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(2.0 * M_PI * u2, &sn, &cs);
        z[k / 3][k % 3] = {rad * cs, rad * sn};
    }
    for (int j = 0; j < 3; j++) {
        // classical Gram-Schmidt applied twice ("twice is enough") keeps the
        // columns orthogonal to rounding even for ill-conditioned draws
        for (int pass = 0; pass < 2; pass++)
            for (int p = 0; p < j; p++) {
                cpx<double> dot = {0.0, 0.0};
                for (int i = 0; i < 3; i++) dot = madd_cl(dot, z[i][p], z[i][j]);
                for (int i = 0; i < 3; i++) z[i][j] = z[i][j] - dot * z[i][p];
            }
        double nn = 0.0;
        for (int i = 0; i < 3; i++) nn += norm2(z[i][j]);
        const double inv = 1.0 / sqrt(nn);
        for (int i = 0; i < 3; i++) z[i][j] = inv * z[i][j];
    }
    const cpx<double> det = z[0][0] * (z[1][1] * z[2][2] - z[1][2] * z[2][1]) -
                            z[0][1] * (z[1][0] * z[2][2] - z[1][2] * z[2][0]) +
                            z[0][2] * (z[1][0] * z[2][1] - z[1][1] * z[2][0]);
    const double ph = -atan2(det.i, det.r) / 3.0;
    double sn, cs;
    sincos(ph, &sn, &cs);
    const cpx<double> rot = {cs, sn};
    mat3<double> m;
    for (int k = 0; k < 9; k++) m.e[k / 3][k % 3] = z[k / 3][k % 3] * rot;
    if constexpr (P == prec::f64)
        store_link64(U, V, mu, s, m);
    else
        store_link32c(U, V, mu, s, m);
}

template <prec P>
__global__ __launch_bounds__(256) void k_cold_start(typename storage<P>::vec* __restrict__ U,
                                                    unsigned V) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= V) return;
    mat3<typename storage<P>::real> m = zero3<typename storage<P>::real>();
    m.e[0][0].r = m.e[1][1].r = m.e[2][2].r = 1;
    if constexpr (P == prec::f64)
        store_link64(U, V, threadIdx.y, s, m);
    else
        store_link32c(U, V, threadIdx.y, s, m);
}

// Global translation U'_mu(x) = U_mu(x - shift) on the periodic lattice, every
// direction and every stored word of the links (mpi::shift::shift_field,
// src/mpi/Shift.cpp:109-160: the reference moves slices between ranks and
// shifts the rest locally; on one device it is a single gather).
template <prec P>
__global__ __launch_bounds__(256) void k_shift(const typename storage<P>::vec* __restrict__ in,
                                               typename storage<P>::vec* __restrict__ out,
                                               lattice g, int4 sh) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= g.V) return;
    unsigned r = s, src = 0;
    const int shv[4] = {sh.x, sh.y, sh.z, sh.w};
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const unsigned L = (unsigned)g.L[d];
        const unsigned c = d < 3 ? r % L : r;
        r = d < 3 ? r / L : 0u;
        src += ((c + L - (unsigned)shv[d]) % L) * g.stride[d];
    }
    constexpr int W = storage<P>::link_words;
    const int mu = threadIdx.y;
#pragma unroll
    for (int k = 0; k < W; k++)
        out[(size_t)(mu * W + k) * g.V + s] = in[(size_t)(mu * W + k) * g.V + src];
}

hipError_t launch_shift(prec p, const void* in, void* out, const lattice& g, const int shift[4],
                        hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const int4 sh = make_int4(shift[0], shift[1], shift[2], shift[3]);
    const dim3 grid((g.V + kSitesPerBlock - 1) / kSitesPerBlock), block(kSitesPerBlock, 4);
    if (p == prec::f64)
        k_shift<prec::f64><<<grid, block, 0, st>>>(static_cast<const double2*>(in),
                                                    static_cast<double2*>(out), g, sh);
    else
        k_shift<prec::f32m><<<grid, block, 0, st>>>(static_cast<const float4*>(in),
                                                     static_cast<float4*>(out), g, sh);
    return hipGetLastError();
}

// Third row of every FP64 link rebuilt from the first two, in place: makes a
// flow output (two stored rows) a complete nine-entry field again.
__global__ __launch_bounds__(256) void k_complete_rows64(double2* __restrict__ U, unsigned V) {
    const unsigned s = blockIdx.x * kSitesPerBlock + threadIdx.x;
    if (s >= V) return;
    double2* b = U + (size_t)threadIdx.y * 9 * V + s;
    cpx<double> m[2][3];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const double2 v = b[(size_t)k * V];
        m[k / 3][k % 3] = {v.x, v.y};
    }
    const cpx<double> r0 = conj(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
    const cpx<double> r1 = conj(m[0][2] * m[1][0] - m[0][0] * m[1][2]);
    const cpx<double> r2 = conj(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
    b[(size_t)6 * V] = make_double2(r0.r, r0.i);
    b[(size_t)7 * V] = make_double2(r1.r, r1.i);
    b[(size_t)8 * V] = make_double2(r2.r, r2.i);
}

hipError_t launch_complete_rows(prec p, void* U, unsigned V, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (p != prec::f64 || !kF64FlowStoresTwoRows) return hipSuccess;
    k_complete_rows64<<<(V + kSitesPerBlock - 1) / kSitesPerBlock, dim3(kSitesPerBlock, 4), 0, st>>>(
        static_cast<double2*>(U), V);
    return hipGetLastError();
}

hipError_t launch_hot_start(prec p, void* U, const lattice& g, uint64_t seed, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const unsigned nb = (g.V + kSitesPerBlock - 1) / kSitesPerBlock;
    if (p == prec::f64)
        k_hot_start<prec::f64><<<nb, dim3(kSitesPerBlock, 4), 0, st>>>(static_cast<double2*>(U),
                                                                        g.V, seed);
    else
        k_hot_start<prec::f32m><<<nb, dim3(kSitesPerBlock, 4), 0, st>>>(static_cast<float4*>(U),
                                                                         g.V, seed);
    return hipGetLastError();
}
hipError_t launch_cold_start(prec p, void* U, const lattice& g, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const unsigned nb = (g.V + kSitesPerBlock - 1) / kSitesPerBlock;
    if (p == prec::f64)
        k_cold_start<prec::f64><<<nb, dim3(kSitesPerBlock, 4), 0, st>>>(static_cast<double2*>(U),
                                                                         g.V);
    else
        k_cold_start<prec::f32m><<<nb, dim3(kSitesPerBlock, 4), 0, st>>>(static_cast<float4*>(U),
                                                                          g.V);
    return hipGetLastError();
}

// ---- exp batch (exp_su3_luscher on independent matrices) --------------------
__global__ __launch_bounds__(256) void k_exp_batch64(const double* __restrict__ Z,
                                                     double* __restrict__ out, int n,
                                                     double coeff, int* __restrict__ warn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* z = Z + (size_t)k * 18;
    alg8<double> x;
    // X = coeff * Z for a traceless anti-hermitian Z (src/su3/utils.cpp:147)
    x.a[0] = coeff * z[1];   // Im Z00
    x.a[1] = coeff * z[9];   // Im Z11
    x.a[2] = coeff * z[2];   // Z01
    x.a[3] = coeff * z[3];
    x.a[4] = coeff * z[4];   // Z02
    x.a[5] = coeff * z[5];
    x.a[6] = coeff * z[10];  // Z12
    x.a[7] = coeff * z[11];
    bool w = false;
    const mat3<double> e = exp_su3<double, double>(x, &w);
    double* o = out + (size_t)k * 18;
    for (int q = 0; q < 9; q++) {
        o[2 * q] = e.e[q / 3][q % 3].r;
        o[2 * q + 1] = e.e[q / 3][q % 3].i;
    }
    if (w) atomicOr(warn, 1);
}

hipError_t launch_exp_batch_f64(const double* Z, double* out, int n, double coeff, int* warn,
                                hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    if (n <= 0) return hipSuccess;
    k_exp_batch64<<<(n + 255) / 256, 256, 0, st>>>(Z, out, n, coeff, warn);
    return hipGetLastError();
}

// ---- plain copy (bandwidth calibration) -------------------------------------
// One 16-byte word per thread on a flat grid: of the copy shapes measured on
// MI355X (grid-stride loops, 4 or 8 words per thread, non-temporal) the
// fastest, 6.2 TB/s read + write on 1-4 GiB.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(const double2* src, double2* dst, size_t n, hipStream_t st) {
    (void)hipGetLastError();  // drop a stale error of an earlier call
    const size_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    k_copy<<<(unsigned)(blocks ? blocks : 1), 256, 0, st>>>(src, dst, n);
    return hipGetLastError();
}

}  // namespace qcb
