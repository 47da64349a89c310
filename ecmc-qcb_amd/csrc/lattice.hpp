// lattice.hpp -- periodic 4-D lattice indexing and the structure-of-arrays
// link layout in HBM.
//
// Site index s = ((t*Lz + z)*Ly + y)*Lx + x (x fastest), as GeometryCB::index
// (src/geometry/GeometryCB.h:25-27) but over the whole periodic lattice, with
// neighbours computed on the fly instead of the reference's 8-entry table plus
// width-1 halos (src/geometry/GeometryCB.cpp:10-31,
// src/mpi/HalosExchange.cpp:4-72).  Extents need not be powers of two.
//
// Link field layout (FP64 path): complex entry k = 3i+j of U_mu(s) lives at
//   U[(mu*9 + k)*V + s]   (one double2 = (re, im) per entry),
// so a wavefront reading 64 consecutive sites issues 16-byte coalesced loads.
// su(3) fields (force / RK3 accumulator) hold 8 reals per link as 4 double2:
//   A[(mu*4 + k)*V + s].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "su3.hpp"

namespace qcb {

struct lattice {
    int L[4];           // Lx, Ly, Lz, Lt
    unsigned V;         // number of sites
    unsigned stride[4]; // 1, Lx, Lx*Ly, Lx*Ly*Lz
    // workgroup -> site-tile order of the flow kernels (see tile_of_block)
    int map_mode;       // 0 linear, 1 one chunk per XCD, 2 XCD chunks per window, 3 slab sweep
    unsigned map_win;   // window (mode 2, workgroups) or slab width (mode 3, tiles)
    // Shape of the 64-site tile of one stage workgroup: lg_tile[d] = log2 of its
    // extent along d (sum 6), or lg_tile[0] < 0 for 64 consecutive sites.  A
    // box tile needs every L[d] divisible by its extent.
    int lg_tile[4];
    unsigned ntile[4];  // tiles along each direction (box tiles)
    // Directions per wavefront of the plain stage kernels (1, 2 or 4; threads
    // per site = 4 / dpw): a workgroup of four wavefronts takes dpw tiles.
    int dpw;
    // FP64 stage kernel with the tile's shared links staged in LDS
    // (stage_lds.hip) where it applies: 1, or 0 for the plain kernel.
    int lds_stage;
};

// Site and coordinates of lane `lane` (0..63) of flow-kernel tile `tile`.
__host__ __device__ inline unsigned tile_site(const lattice& g, unsigned tile, unsigned lane,
                                     unsigned (&c)[4]) {
    if (g.lg_tile[0] < 0) {
        const unsigned s = tile * 64u + lane;
        unsigned r = s;
        c[0] = r % (unsigned)g.L[0];
        r /= (unsigned)g.L[0];
        c[1] = r % (unsigned)g.L[1];
        r /= (unsigned)g.L[1];
        c[2] = r % (unsigned)g.L[2];
        c[3] = r / (unsigned)g.L[2];
        return s;
    }
    unsigned r = tile, l = lane;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const unsigned td = (d < 3) ? r % g.ntile[d] : r;
        r = (d < 3) ? r / g.ntile[d] : 0u;
        const unsigned ld = l & ((1u << g.lg_tile[d]) - 1u);
        l >>= g.lg_tile[d];
        c[d] = (td << g.lg_tile[d]) + ld;
    }
    return ((c[3] * (unsigned)g.L[2] + c[2]) * (unsigned)g.L[1] + c[1]) * (unsigned)g.L[0] + c[0];
}

// Tiles of one t-slice for the slab order (0 when 64-site tiles straddle
// t-slices): stage tiles (64 consecutive sites or a box) and the measurement's
// 64 consecutive sites.
__host__ __device__ inline unsigned stage_tiles_per_t(const lattice& g) {
    if (g.lg_tile[0] >= 0) return g.ntile[0] * g.ntile[1] * g.ntile[2];
    const unsigned n = (unsigned)g.L[0] * (unsigned)g.L[1] * (unsigned)g.L[2];
    return n % 64u == 0u ? n / 64u : 0u;
}
__host__ __device__ inline unsigned site_tiles_per_t(const lattice& g) {
    const unsigned n = (unsigned)g.L[0] * (unsigned)g.L[1] * (unsigned)g.L[2];
    return n % 64u == 0u ? n / 64u : 0u;
}

// The dispatcher hands workgroups to the 8 XCDs round-robin (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  These orders give every XCD a
// contiguous run of site tiles so that neighbouring tiles share its L2:
//   mode 1: XCD k takes tiles [k q, (k+1) q), q = ceil(ntiles / 8);
//   mode 2: the same inside consecutive windows of win workgroups (win rounded
//           down to a multiple of 8), so all XCDs still sweep the lattice
//           together (t-neighbours stay in the shared last-level cache);
//   mode 3: slab sweep.  The per_t tiles of one t-slice are cut into slabs of
//           win tiles (slab = a run of z-slices when win is a multiple of the
//           tiles of one z-slice); XCD k owns slabs k, k+8, ... and walks each
//           through t, t-slice by t-slice.  The tiles at t-1, t and t+1 of a
//           slab are then at most 2 win dispatches apart on the same XCD, so
//           the t-neighbours of a link are still in that XCD's L2 when they
//           are read again.  Needs per_t > 0 (tiles never straddle t-slices);
//           otherwise mode 3 is the linear order.
// Returns a tile index >= ntiles for padding workgroups, which exit.
__host__ __device__ inline unsigned tile_of_block(unsigned b, unsigned ntiles, int mode,
                                                  unsigned win, unsigned per_t = 0u) {
    if (mode == 1) {
        const unsigned q = (ntiles + 7) / 8;
        return (b % 8) * q + b / 8;
    }
    if (mode == 2) {
        const unsigned w8 = win < 8u ? 8u : (win & ~7u);
        const unsigned w = b / w8, j = b % w8;
        return w * w8 + (j % 8) * (w8 / 8) + j / 8;
    }
    if (mode == 3 && per_t > 0u) {
        const unsigned R = win < 1u ? 1u : win;
        const unsigned nt = ntiles / per_t;     // t-slices of tiles
        const unsigned k = b / 8, x = b % 8;    // k-th workgroup of XCD x
        const unsigned slab = 8 * (k / (nt * R)) + x;
        const unsigned p = k % (nt * R);
        const unsigned r = slab * R + p % R;    // tile within the t-slice
        return r < per_t ? (p / R) * per_t + r : ntiles;
    }
    return b;
}
__host__ __device__ inline unsigned grid_for_tiles(unsigned ntiles, int mode, unsigned win,
                                                   unsigned per_t = 0u) {
    if (mode == 1) return 8 * ((ntiles + 7) / 8);
    if (mode == 2) {
        const unsigned w8 = win < 8u ? 8u : (win & ~7u);
        return ((ntiles + w8 - 1) / w8) * w8;
    }
    if (mode == 3 && per_t > 0u) {
        const unsigned R = win < 1u ? 1u : win;
        const unsigned slabs = (per_t + R - 1) / R;
        return 8 * ((slabs + 7) / 8) * (ntiles / per_t) * R;
    }
    return ntiles;
}

inline lattice make_lattice(const int dims[4]) {
    lattice g;
    unsigned st = 1;
    for (int d = 0; d < 4; d++) {
        g.L[d] = dims[d];
        g.stride[d] = st;
        st *= (unsigned)dims[d];
    }
    g.V = st;
    for (int d = 0; d < 4; d++) {
        g.lg_tile[d] = d == 0 ? -1 : 0;
        g.ntile[d] = 0;
    }
    g.dpw = 1;
    g.lds_stage = 1;
    // Slab sweep with 16 slabs per t-slice, or as many more (32 ... 512) as
    // keep a slab at most 256 tiles: at 96^4 a 16-slab slab is 864 tiles,
    // whose links at t-1, t and t+1 outgrow the caches, and 216-tile slabs ran
    // 9% faster (FP64).  Equal slabs first; then slabs of at most 256 tiles
    // with a shorter last one (the order pads it); then 16 or 8 slabs of any
    // size; else XCD chunks in windows of 1024 (measured on MI355X at 32^4,
    // 48^4, 64^4 and 96^4; qcbf_autotune refines it).
    const unsigned pt = site_tiles_per_t(g);
    g.map_mode = 2;
    g.map_win = 1024;
    for (unsigned ns : {16u, 32u, 64u, 128u, 256u, 512u})
        if (pt && pt % ns == 0 && pt / ns <= 256u) {
            g.map_mode = 3;
            g.map_win = pt / ns;
            break;
        }
    if (g.map_mode == 2 && pt > 16u * 256u)
        for (unsigned ns : {32u, 64u, 128u, 256u, 512u})
            if ((pt + ns - 1) / ns <= 256u) {
                g.map_mode = 3;
                g.map_win = (pt + ns - 1) / ns;
                break;
            }
    if (g.map_mode == 2 && pt && (pt % 16 == 0 || pt % 8 == 0)) {
        g.map_mode = 3;
        g.map_win = pt % 16 == 0 ? pt / 16 : pt / 8;
    }
    return g;
}

// Signed index offsets of the eight nearest neighbours of one site.
struct nbr {
    int up[4];
    int dn[4];
    // Lanes holding the sites x+1 and x-1 when the wavefront's 64 sites are
    // whole x-rows (set_x_lanes); -1 otherwise.
    int xp = -1, xm = -1;
};

// A wavefront of a 64-consecutive-site tile holds whole x-rows when Lx divides
// 64 and every tile is full: the links at x+1 and x-1 are then in the same
// wavefront, and the staple sum can take them from a lane shuffle instead of
// a load (staple_sum, flow_body.hpp).  All 64 lanes must take part.
__host__ __device__ inline bool x_rows_in_wave(const lattice& g) {
    return g.lg_tile[0] < 0 && g.L[0] <= 64 && 64 % g.L[0] == 0 && g.V % 64 == 0;
}
__device__ inline void set_x_lanes(nbr& n, const lattice& g, unsigned lane) {
    const unsigned Lx = (unsigned)g.L[0], x = lane & (Lx - 1), base = lane - x;
    n.xp = (int)(base + ((x + 1) & (Lx - 1)));
    n.xm = (int)(base + ((x + Lx - 1) & (Lx - 1)));
}

__device__ inline nbr neighbours_at(const lattice& g, const unsigned (&c)[4]) {
    nbr n;
#pragma unroll
    for (int d = 0; d < 4; d++) {
        const int st = (int)g.stride[d];
        const int wrap = (g.L[d] - 1) * st;
        n.up[d] = (c[d] == (unsigned)(g.L[d] - 1)) ? -wrap : st;
        n.dn[d] = (c[d] == 0u) ? wrap : -st;
    }
    return n;
}

__device__ inline nbr neighbours(const lattice& g, unsigned s) {
    unsigned c[4];
    unsigned r = s;
    c[0] = r % (unsigned)g.L[0];
    r /= (unsigned)g.L[0];
    c[1] = r % (unsigned)g.L[1];
    r /= (unsigned)g.L[1];
    c[2] = r % (unsigned)g.L[2];
    c[3] = r / (unsigned)g.L[2];
    return neighbours_at(g, c);
}

// Set the stage-tile shape (log2 extents summing to 6) if every extent is
// divisible by it; lg = nullptr (or a shape that does not divide the lattice)
// selects 64 consecutive sites.  Returns whether the box shape was taken.
inline bool set_tile_shape(lattice& g, const int* lg) {
    bool ok = lg != nullptr;
    if (ok) {
        int sum = 0;
        for (int d = 0; d < 4; d++) {
            ok = ok && lg[d] >= 0 && lg[d] <= 6 && g.L[d] % (1 << lg[d]) == 0;
            sum += lg[d];
        }
        ok = ok && sum == 6;
    }
    for (int d = 0; d < 4; d++) {
        g.lg_tile[d] = ok ? lg[d] : (d == 0 ? -1 : 0);
        g.ntile[d] = ok ? (unsigned)(g.L[d] >> lg[d]) : 0u;
    }
    return ok;
}
__host__ __device__ inline unsigned stage_tiles(const lattice& g) {
    return g.lg_tile[0] < 0 ? (g.V + 63) / 64 : g.ntile[0] * g.ntile[1] * g.ntile[2] * g.ntile[3];
}

// A plain-stage workgroup with dpw directions per wavefront covers a group of
// dpw consecutive tiles (4/dpw wavefronts each).  The XCD-aware order then
// maps tile groups: n groups, per_t groups per t-slice (0 when a group would
// straddle two t-slices, which falls back to the linear order in mode 3), and
// a mode-3 slab of win/dpw groups, the same number of sites as win tiles.
struct tile_groups {
    unsigned n, per_t, win;
};
__host__ __device__ inline tile_groups group_tiles(unsigned ntiles, unsigned per_t, int mode,
                                                   unsigned win, unsigned dpw) {
    tile_groups r;
    r.n = (ntiles + dpw - 1) / dpw;
    r.per_t = per_t % dpw == 0 ? per_t / dpw : 0u;
    r.win = mode == 3 ? (win / dpw < 1u ? 1u : win / dpw) : win;
    return r;
}
// Tile of wavefront y (the row threadIdx.y of a 64 x 4 block) in tile group
// grp: 4/dpw wavefronts per tile.
__host__ __device__ inline unsigned tile_in_group(unsigned grp, unsigned y, unsigned dpw) {
    return grp * dpw + y / (4u / dpw);
}
// A wavefront is one row y of the block, so y is wavefront-uniform;
// readfirstlane tells the compiler, which keeps the tile and what is read
// through it (the ensemble descriptors) in scalar registers.
template <int DPW>
__device__ inline unsigned group_tile(unsigned grp, unsigned y) {
    if constexpr (DPW == 1) return grp;
    return tile_in_group(grp, (unsigned)__builtin_amdgcn_readfirstlane((int)y), DPW);
}
__host__ __device__ inline tile_groups stage_groups(const lattice& g) {
    return group_tiles(stage_tiles(g), stage_tiles_per_t(g), g.map_mode, g.map_win,
                       (unsigned)g.dpw);
}

// Word s of a word array of a field (16-byte words).  The array's base is the
// same for the whole wavefront (scalar registers) and the site enters as a
// 32-bit byte offset, so each load or store takes both as they are instead of
// a 64-bit address formed per access.  Needs 16 s < 2^32: fields of at most
// 2^28 sites (kMaxSites32c; qcbf_create and qcbf_batch_create take volumes
// below it).
constexpr unsigned kMaxSites32c = 1u << 28;
template <class W>
__device__ inline W* word_ptr(W* base, unsigned s) {
    static_assert(sizeof(W) == 16, "16-byte words");
    using B = typename std::conditional<std::is_const<W>::value, const char, char>::type;
    return reinterpret_cast<W*>(reinterpret_cast<B*>(base) + (s << 4));
}

// ---- FP64 structure-of-arrays link field -----------------------------------
// The flow and measurement kernels read only the first two rows of every
// FP64 link and rebuild the third as conj(row0 x row1), exact for SU(3) up to
// rounding: a third fewer cache lines per link, which is what bounds these
// kernels on MI355X (the field in HBM keeps all nine entries).  With
// QCB_F64_RECONSTRUCT=0 all nine are read, and one operand of every staple
// product is streamed a row at a time (QCB_F64_STREAM_ROWS).
#ifndef QCB_F64_STREAM_ROWS
#define QCB_F64_STREAM_ROWS 1
#endif
#ifndef QCB_F64_RECONSTRUCT
#define QCB_F64_RECONSTRUCT 1
#endif
struct field64 {
    using real = double;
    using word = double2;
    static constexpr bool stream_rows = QCB_F64_STREAM_ROWS && !QCB_F64_RECONSTRUCT;
    // words read per link: two rows (third rebuilt) or all nine entries
    static constexpr int kWords = QCB_F64_RECONSTRUCT ? 6 : 9;
    const double2* __restrict__ p;
    unsigned V;
    __device__ inline void load_words(int mu, unsigned s, word (&w)[kWords]) const {
#pragma unroll
        for (int k = 0; k < kWords; k++) w[k] = *word_ptr(p + ((size_t)mu * 9 + k) * V, s);
    }
    __device__ static inline mat3<double> from_words(const word (&w)[kWords]) {
        mat3<double> m;
#pragma unroll
        for (int k = 0; k < kWords; k++) m.e[k / 3][k % 3] = {w[k].x, w[k].y};
#if QCB_F64_RECONSTRUCT
        // third row of an SU(3) matrix: conj(row0 x row1)
        m.e[2][0] = conj(m.e[0][1] * m.e[1][2] - m.e[0][2] * m.e[1][1]);
        m.e[2][1] = conj(m.e[0][2] * m.e[1][0] - m.e[0][0] * m.e[1][2]);
        m.e[2][2] = conj(m.e[0][0] * m.e[1][1] - m.e[0][1] * m.e[1][0]);
#endif
        return m;
    }
    __device__ inline mat3<double> link(int mu, unsigned s) const {
        word w[kWords];
        load_words(mu, s, w);
        return from_words(w);
    }
    // row i of U_mu(s)
    __device__ inline void row(int mu, unsigned s, int i, cpx<double> (&r)[3]) const {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double2 v = *word_ptr(p + ((size_t)mu * 9 + 3 * i + j) * V, s);
            r[j] = {v.x, v.y};
        }
    }
};

#ifndef QCB_NT_STORES
#define QCB_NT_STORES 1
#endif
// Stores of the flow's output fields are non-temporal: the next stage reads
// them back from HBM anyway, and keeping them out of L2 leaves it to the
// neighbour links the current stage re-reads.
// One 16-byte store per lane (a vector store never splits into narrower ones).
typedef double double2_vec __attribute__((ext_vector_type(2)));
__device__ inline void put2(double2* p, double2 v) {
#if QCB_NT_STORES
    const double2_vec w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<double2_vec*>(p));
#else
    *p = v;
#endif
}

__device__ inline void store_link64(double2* __restrict__ p, unsigned V, int mu, unsigned s,
                                    const mat3<double>& m) {
#pragma unroll
    for (int k = 0; k < 9; k++)
        put2(word_ptr(p + ((size_t)mu * 9 + k) * V, s),
             make_double2(m.e[k / 3][k % 3].r, m.e[k / 3][k % 3].i));
}

// The flow kernels store rows 0 and 1 only when every reader rebuilds the
// third (QCB_F64_RECONSTRUCT): a third less write traffic per stage.  The
// host completes the third row (launch_complete_rows) before the field
// leaves the device.
constexpr bool kF64FlowStoresTwoRows = QCB_F64_RECONSTRUCT != 0;
__device__ inline void store_flow_link64(double2* __restrict__ p, unsigned V, int mu, unsigned s,
                                         const mat3<double>& m) {
#pragma unroll
    for (int k = 0; k < (kF64FlowStoresTwoRows ? 6 : 9); k++)
        put2(word_ptr(p + ((size_t)mu * 9 + k) * V, s),
             make_double2(m.e[k / 3][k % 3].r, m.e[k / 3][k % 3].i));
}

// The RK3 accumulator is read once per stage and never by a neighbour:
// non-temporal loads keep it out of L2, which holds the neighbour links the
// stage re-reads (measured 1.5-2% on the FP64 stage kernels at 32^4 and 64^4).
#ifndef QCB_NT_ACC_LOADS
#define QCB_NT_ACC_LOADS 1
#endif
__device__ inline alg8<double> load_alg64(const double2* __restrict__ p, unsigned V, int mu,
                                          unsigned s) {
    alg8<double> a;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double2* w = word_ptr(p + ((size_t)mu * 4 + k) * V, s);
#if QCB_NT_ACC_LOADS
        a.a[2 * k] = __builtin_nontemporal_load(&w->x);
        a.a[2 * k + 1] = __builtin_nontemporal_load(&w->y);
#else
        const double2 v = *w;
        a.a[2 * k] = v.x;
        a.a[2 * k + 1] = v.y;
#endif
    }
    return a;
}

__device__ inline void store_alg64(double2* __restrict__ p, unsigned V, int mu, unsigned s,
                                   const alg8<double>& a) {
#pragma unroll
    for (int k = 0; k < 4; k++)
        put2(word_ptr(p + ((size_t)mu * 4 + k) * V, s), make_double2(a.a[2 * k], a.a[2 * k + 1]));
}

}  // namespace qcb
