// qcbflow_batch.cpp -- C ABI of the ensemble path (include/qcbflow.h,
// "ensembles" section).
//
// An ensemble of lattices of uneven volumes is flowed together: one persistent
// kernel per RK3 stage over all configurations, per-configuration
// measurements in one measurement + one reduction launch, and, optionally,
// every (measure, RK3 step) pair captured once in a hipGraph and replayed.  The
// reference has no ensemble mode: its callers flow one configuration at a time
// through topo_charge_flowed (src/observables/observables_mpi.cpp:159-186);
// the per-configuration series produced here follow that cadence exactly.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "batch.hpp"
#include "kernels.hpp"
#include "qcbflow.h"

using qcb::cfg_desc;

struct qcbf_batch {
    int dev = 0;
    int precision = QCBF_FP64;
    qcb::prec prec = qcb::prec::f64;
    int n = 0;
    std::vector<cfg_desc> cfgs;  // host copy
    unsigned ntiles = 0;
    size_t link_words = 0, alg_words = 0;
    hipStream_t stream = nullptr;
    char* U[2] = {nullptr, nullptr};
    int cur = 0;
    std::vector<char> rows_stale;  // per configuration: field holds rows 0-1 only
    char* acc = nullptr;
    cfg_desc* d_cfgs = nullptr;
    unsigned* d_tile_cfg = nullptr;
    double* partials = nullptr;
    double* dres = nullptr;
    size_t dres_cap = 0;
    unsigned* d_slot = nullptr;
    int* dwarn = nullptr;
    void* stage = nullptr;
    qcb::batch_launch bl{0, 2, 1024, 1, 0};
    int lds_stage = 1;       // QCBF_TUNE_LDS_STAGE
    bool full_tiles = true;  // every member's volume a multiple of 64
    int blocks_per_cu = 0;  // one workgroup per tile: the autotuner's pick on MI355X
    long tile_shape = -1;
    int cus = 256;
    bool use_graph = false;
    bool fuse_measure = true;  // flow_measure: measurement inside the next stage 1
    // graph of one (measure, RK3 step) pair, keyed by (cur, eps)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double geps[2] = {0.0, 0.0};
    std::string err;
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double step_ms = 0.0;
    long step_n = 0;
};

static int bfail(qcbf_batch* b, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

static thread_local std::string g_batch_err;

#define BCHK(b, expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (void)hipGetLastError();                                                       \
            return bfail((b), e_ == hipErrorOutOfMemory ? QCBF_ERR_OOM : QCBF_ERR_HIP,     \
                         "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
        }                                                                                  \
    } while (0)

#define CHECK_B(b)                                                              \
    do {                                                                        \
        if (!(b)) {                                                             \
            g_batch_err = "null ensemble handle";                               \
            return QCBF_ERR_ARG;                                                \
        }                                                                       \
        BCHK((b), hipSetDevice((b)->dev));                                      \
    } while (0)

constexpr unsigned kChunkSites = 65536;
constexpr size_t kChunkBytes = (size_t)kChunkSites * 72 * sizeof(double);

static void drop_graphs(qcbf_batch* b) {
    for (int k = 0; k < 2; k++)
        if (b->gexec[k]) {
            (void)hipGraphExecDestroy(b->gexec[k]);
            b->gexec[k] = nullptr;
        }
}

static void release(qcbf_batch* b) {
    (void)hipSetDevice(b->dev);
    drop_graphs(b);
    for (auto* p : {b->U[0], b->U[1], b->acc}) (void)hipFree(p);
    (void)hipFree(b->d_cfgs);
    (void)hipFree(b->d_tile_cfg);
    (void)hipFree(b->partials);
    (void)hipFree(b->dres);
    (void)hipFree(b->d_slot);
    (void)hipFree(b->dwarn);
    (void)hipFree(b->stage);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" const char* qcbf_batch_last_error(const qcbf_batch* b) {
    return b ? b->err.c_str() : g_batch_err.c_str();
}

static void set_grid(qcbf_batch* b) {
    b->bl.grid = (unsigned)(b->cus * b->blocks_per_cu);
    b->bl.lds = b->lds_stage && b->full_tiles && b->tile_shape == -1 && b->bl.grid == 0 &&
                b->bl.dpw == 1;
    drop_graphs(b);
}

extern "C" int qcbf_batch_create(int n, const int* dims, int precision, qcbf_batch** out) {
    if (!out || !dims || n <= 0) {
        g_batch_err = "bad arguments";
        return QCBF_ERR_ARG;
    }
    *out = nullptr;
    if (precision != QCBF_FP64 && precision != QCBF_FP32_MIXED) {
        g_batch_err = "unknown precision";
        return QCBF_ERR_ARG;
    }
    qcbf_batch* b = new qcbf_batch();
    b->n = n;
    b->rows_stale.assign(n, 0);
    b->precision = precision;
    b->prec = precision == QCBF_FP64 ? qcb::prec::f64 : qcb::prec::f32m;
    const size_t lw_per_site = precision == QCBF_FP64 ? 36 : 12;  // link-field words per site
    const size_t aw_per_site = precision == QCBF_FP64 ? 16 : 8;   // su(3)-field words per site
    size_t uo = 0, ao = 0;
    unsigned t0 = 0;
    std::vector<unsigned> tile_cfg;
    for (int c = 0; c < n; c++) {
        unsigned long long V = 1;
        for (int d = 0; d < 4; d++) {
            const int L = dims[4 * c + d];
            if (L < 2 || L > 4096) {
                delete b;
                g_batch_err = "extent out of range [2, 4096] in configuration " + std::to_string(c);
                return QCBF_ERR_ARG;
            }
            V *= (unsigned long long)L;
        }
        static_assert((1ull << 28) <= qcb::kMaxSites32c, "volume limit beyond 32-bit site offsets");
        if (V >= (1ull << 28)) {
            delete b;
            g_batch_err = "configuration volume too large";
            return QCBF_ERR_ARG;
        }
        cfg_desc cd;
        cd.g = qcb::make_lattice(&dims[4 * c]);
        cd.tile0 = t0;
        cd.ntiles = (cd.g.V + qcb::kSitesPerBlock - 1) / qcb::kSitesPerBlock;
        if (cd.g.V % qcb::kSitesPerBlock != 0) b->full_tiles = false;
        cd.u_off = uo;
        cd.a_off = ao;
        for (unsigned k = 0; k < cd.ntiles; k++) tile_cfg.push_back((unsigned)c);
        t0 += cd.ntiles;
        // keep every field 256-byte aligned
        uo += ((size_t)cd.g.V * lw_per_site + 15) / 16 * 16;
        ao += ((size_t)cd.g.V * aw_per_site + 15) / 16 * 16;
        b->cfgs.push_back(cd);
    }
    b->ntiles = t0;
    b->link_words = uo;
    b->alg_words = ao;
    const size_t wbytes = precision == QCBF_FP64 ? sizeof(double2) : sizeof(float4);
    auto fail_create = [&](hipError_t e, const char* what) {
        (void)hipGetLastError();
        g_batch_err = std::string(what) + ": " + hipGetErrorString(e);
        const int code = e == hipErrorOutOfMemory ? QCBF_ERR_OOM : QCBF_ERR_HIP;
        release(b);
        return code;
    };
    hipError_t e = hipGetDevice(&b->dev);
    if (e != hipSuccess) return fail_create(e, "hipGetDevice");
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, b->dev)) != hipSuccess)
        return fail_create(e, "hipGetDeviceProperties");
    b->cus = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail_create(e, "hipStreamCreate");
    for (int k = 0; k < 2; k++)
        if ((e = hipMalloc(&b->U[k], uo * wbytes)) != hipSuccess) return fail_create(e, "hipMalloc");
    if ((e = hipMalloc(&b->acc, ao * wbytes)) != hipSuccess) return fail_create(e, "hipMalloc");
    if ((e = hipMalloc(&b->d_cfgs, sizeof(cfg_desc) * n)) != hipSuccess)
        return fail_create(e, "hipMalloc");
    if ((e = hipMalloc(&b->d_tile_cfg, sizeof(unsigned) * t0)) != hipSuccess)
        return fail_create(e, "hipMalloc");
    if ((e = hipMalloc(&b->partials, sizeof(double) * 3 * t0)) != hipSuccess)
        return fail_create(e, "hipMalloc");
    if ((e = hipMalloc(&b->d_slot, sizeof(unsigned))) != hipSuccess) return fail_create(e, "hipMalloc");
    if ((e = hipMalloc(&b->dwarn, sizeof(int))) != hipSuccess) return fail_create(e, "hipMalloc");
    b->dres_cap = 3 * (size_t)n;
    if ((e = hipMalloc(&b->dres, sizeof(double) * b->dres_cap)) != hipSuccess)
        return fail_create(e, "hipMalloc");
    if ((e = hipMemcpy(b->d_cfgs, b->cfgs.data(), sizeof(cfg_desc) * n, hipMemcpyHostToDevice)) !=
        hipSuccess)
        return fail_create(e, "hipMemcpy");
    if ((e = hipMemcpy(b->d_tile_cfg, tile_cfg.data(), sizeof(unsigned) * t0,
                       hipMemcpyHostToDevice)) != hipSuccess)
        return fail_create(e, "hipMemcpy");
    if ((e = hipMemset(b->dwarn, 0, sizeof(int))) != hipSuccess) return fail_create(e, "hipMemset");
    for (int c = 0; c < n; c++) {
        const cfg_desc& cd = b->cfgs[c];
        if ((e = qcb::launch_cold_start(b->prec, b->U[0] + cd.u_off * wbytes, cd.g, b->stream)) !=
            hipSuccess)
            return fail_create(e, "cold start");
    }
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail_create(e, "sync");
    set_grid(b);
    *out = b;
    return QCBF_OK;
}

extern "C" int qcbf_batch_destroy(qcbf_batch* b) {
    if (!b) return QCBF_OK;
    (void)hipSetDevice(b->dev);
    (void)hipStreamSynchronize(b->stream);
    release(b);
    return QCBF_OK;
}

extern "C" int qcbf_batch_size(const qcbf_batch* b, int* n, int* dims) {
    if (!b || !n) return QCBF_ERR_ARG;
    *n = b->n;
    if (dims)
        for (int c = 0; c < b->n; c++)
            for (int d = 0; d < 4; d++) dims[4 * c + d] = b->cfgs[c].g.L[d];
    return QCBF_OK;
}

static size_t wbytes(const qcbf_batch* b) {
    return b->precision == QCBF_FP64 ? sizeof(double2) : sizeof(float4);
}
static void* field_of(qcbf_batch* b, int c) { return b->U[b->cur] + b->cfgs[c].u_off * wbytes(b); }

static int check_cfg(qcbf_batch* b, int c) {
    if (c < 0 || c >= b->n) return bfail(b, QCBF_ERR_ARG, "configuration %d out of range", c);
    return QCBF_OK;
}

static int ensure_stage(qcbf_batch* b) {
    if (!b->stage) BCHK(b, hipMalloc(&b->stage, kChunkBytes));
    return QCBF_OK;
}

extern "C" int qcbf_batch_set_links(qcbf_batch* b, int c, const double* links) {
    CHECK_B(b);
    int rc = check_cfg(b, c);
    if (rc) return rc;
    if (!links) return bfail(b, QCBF_ERR_ARG, "null links");
    if ((rc = ensure_stage(b))) return rc;
    const qcb::lattice& g = b->cfgs[c].g;
    for (unsigned s0 = 0; s0 < g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, g.V - s0);
        BCHK(b, hipMemcpyAsync(b->stage, links + (size_t)s0 * 72, (size_t)n * 72 * sizeof(double),
                               hipMemcpyHostToDevice, b->stream));
        BCHK(b, qcb::launch_aos_to_field(b->prec, b->stage, 0, field_of(b, c), g, s0, n, b->stream));
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    b->rows_stale[c] = 0;
    return QCBF_OK;
}

// Flow steps store FP64 links as two rows; mark every configuration.
static void mark_flowed(qcbf_batch* b) {
    const bool two = b->prec == qcb::prec::f64 && qcb::kF64FlowStoresTwoRows;
    for (auto& f : b->rows_stale) f = two ? 1 : 0;
}

extern "C" int qcbf_batch_get_links(qcbf_batch* b, int c, double* links) {
    CHECK_B(b);
    int rc = check_cfg(b, c);
    if (rc) return rc;
    if (!links) return bfail(b, QCBF_ERR_ARG, "null links");
    if ((rc = ensure_stage(b))) return rc;
    const qcb::lattice& g = b->cfgs[c].g;
    if (b->rows_stale[c]) {
        BCHK(b, qcb::launch_complete_rows(b->prec, field_of(b, c), g.V, b->stream));
        b->rows_stale[c] = 0;
    }
    double* st = static_cast<double*>(b->stage);
    for (unsigned s0 = 0; s0 < g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, g.V - s0);
        BCHK(b, qcb::launch_field_to_aos(b->prec, field_of(b, c), st, g, s0, n, false, b->stream));
        BCHK(b, hipMemcpyAsync(links + (size_t)s0 * 72, st, (size_t)n * 72 * sizeof(double),
                               hipMemcpyDeviceToHost, b->stream));
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    return QCBF_OK;
}

extern "C" int qcbf_batch_hot_start(qcbf_batch* b, int c, uint64_t seed) {
    CHECK_B(b);
    int rc = check_cfg(b, c);
    if (rc) return rc;
    BCHK(b, qcb::launch_hot_start(b->prec, field_of(b, c), b->cfgs[c].g, seed, b->stream));
    b->rows_stale[c] = 0;
    return QCBF_OK;
}

extern "C" int qcbf_batch_copy_from(qcbf_batch* b, int c, qcbf_flow* src) {
    CHECK_B(b);
    int rc = check_cfg(b, c);
    if (rc) return rc;
    int sp = 0;
    const void* sf = nullptr;
    hipStream_t ss = nullptr;
    if (!src || qcb::field_view(src, &sp, &sf, &ss) != QCBF_OK)
        return bfail(b, QCBF_ERR_ARG, "bad source handle");
    int sdims[4], sprec = 0;
    qcbf_get_dims(src, sdims, &sprec);
    const qcb::lattice& g = b->cfgs[c].g;
    for (int d = 0; d < 4; d++)
        if (sdims[d] != g.L[d]) return bfail(b, QCBF_ERR_ARG, "extents differ");
    if ((rc = ensure_stage(b))) return rc;
    BCHK(b, hipStreamSynchronize(ss));
    double* st = static_cast<double*>(b->stage);
    for (unsigned s0 = 0; s0 < g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, g.V - s0);
        BCHK(b, qcb::launch_field_to_aos((qcb::prec)sp, sf, st, g, s0, n, false, b->stream));
        BCHK(b, qcb::launch_aos_to_field(b->prec, st, 0, field_of(b, c), g, s0, n, b->stream));
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    b->rows_stale[c] = 0;
    return QCBF_OK;
}

// ---- flow -----------------------------------------------------------------------
// One RK3 step of every member.  With out set, stage 1 also measures its input
// into result slot *slot (bumped afterwards when bump is set), as
// enqueue_measure would.
static int enqueue_step(qcbf_batch* b, double eps, int cur, double* out = nullptr,
                        unsigned* slot = nullptr, bool bump = false) {
    const double cz[3] = {0.25 * eps, (8.0 / 9.0) * eps, 0.75 * eps};
    const double ca[3] = {0.0, (17.0 / 36.0) * eps, eps};
    for (int stage = 1; stage <= 3; stage++) {
        if (stage == 1 && out)
            BCHK(b, qcb::launch_batch_stage1_measure(b->prec, b->U[cur], b->U[cur ^ 1], b->acc,
                                                     b->d_cfgs, b->d_tile_cfg, b->ntiles, b->n,
                                                     b->bl, cz[0], b->dwarn, b->partials, out,
                                                     slot, bump, b->stream));
        else
            BCHK(b, qcb::launch_batch_stage(b->prec, stage, b->U[cur], b->U[cur ^ 1], b->acc,
                                            b->d_cfgs, b->d_tile_cfg, b->ntiles, b->bl,
                                            cz[stage - 1], ca[stage - 1], b->dwarn, b->stream));
        cur ^= 1;
    }
    return QCBF_OK;
}

// The measurement can ride on the next stage 1 when every member uses
// 64-consecutive-site tiles.
static bool fusable(const qcbf_batch* b) {
    return b->fuse_measure && qcb::kFusedMeasureBuilt && b->tile_shape == -1;
}

static int enqueue_measure(qcbf_batch* b, double* out, unsigned* slot, bool bump) {
    BCHK(b, qcb::launch_batch_measure(b->prec, b->U[b->cur], b->d_cfgs, b->d_tile_cfg, b->ntiles,
                                      b->n, b->bl, b->partials, out, slot, bump, b->stream));
    return QCBF_OK;
}

extern "C" int qcbf_batch_rk3_step(qcbf_batch* b, double eps, int n_steps) {
    CHECK_B(b);
    if (n_steps < 0) return bfail(b, QCBF_ERR_ARG, "n_steps < 0");
    if (n_steps > 0) mark_flowed(b);
    for (int k = 0; k < n_steps; k++) {
        int rc = enqueue_step(b, eps, b->cur);
        if (rc) return rc;
        b->cur ^= 1;  // three stages flip the ping-pong an odd number of times
    }
    return QCBF_OK;
}

extern "C" int qcbf_batch_measure(qcbf_batch* b, double* out) {
    CHECK_B(b);
    if (!out) return bfail(b, QCBF_ERR_ARG, "null output");
    int rc = enqueue_measure(b, b->dres, nullptr, false);
    if (rc) return rc;
    BCHK(b, hipMemcpyAsync(out, b->dres, sizeof(double) * 3 * b->n, hipMemcpyDeviceToHost,
                           b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return QCBF_OK;
}

// One (measurement into slot *d_slot, slot++, RK3 step) unit captured as a graph
// for the current ping-pong parity; rebuilt when eps or the launch shape changes.
static int graph_for(qcbf_batch* b, double eps, hipGraphExec_t* exec) {
    const int cur = b->cur;
    if (b->gexec[cur] && b->geps[cur] == eps) {
        *exec = b->gexec[cur];
        return QCBF_OK;
    }
    if (b->gexec[cur]) {
        (void)hipGraphExecDestroy(b->gexec[cur]);
        b->gexec[cur] = nullptr;
    }
    hipGraph_t graph = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = QCBF_OK;
    if (fusable(b)) {
        rc = enqueue_step(b, eps, cur, b->dres, b->d_slot, true);
    } else {
        rc = enqueue_measure(b, b->dres, b->d_slot, true);
        if (!rc) rc = enqueue_step(b, eps, cur);
    }
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, QCBF_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->gexec[cur], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return bfail(b, QCBF_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    b->geps[cur] = eps;
    *exec = b->gexec[cur];
    return QCBF_OK;
}

extern "C" int qcbf_batch_flow_measure(qcbf_batch* b, double eps, int n_steps_gf, int n_rk_steps,
                                       double* tQE, double* plaq) {
    CHECK_B(b);
    if (n_steps_gf < 0 || n_rk_steps < 0 || (n_steps_gf > 0 && !tQE))
        return bfail(b, QCBF_ERR_ARG, "bad flow_measure arguments");
    if (n_steps_gf == 0) return QCBF_OK;
    const size_t need = (size_t)n_steps_gf * 3 * b->n;
    if (need > b->dres_cap) {
        BCHK(b, hipStreamSynchronize(b->stream));
        (void)hipFree(b->dres);
        b->dres = nullptr;
        b->dres_cap = 0;
        BCHK(b, hipMalloc(&b->dres, need * sizeof(double)));
        b->dres_cap = need;
        drop_graphs(b);  // graphs hold the old results pointer
    }
    if (n_steps_gf > 1 && n_rk_steps > 0) mark_flowed(b);
    if (b->timing) BCHK(b, hipEventRecord(b->ev[0], b->stream));
    BCHK(b, qcb::launch_set_counter(b->d_slot, 0, b->stream));
    std::vector<double> tvals(n_steps_gf);
    double t = 0.0;
    for (int n = 0; n < n_steps_gf; n++) {
        tvals[n] = t;
        const bool last = n == n_steps_gf - 1;
        if (b->use_graph && n_rk_steps == 1 && !last) {
            hipGraphExec_t ex = nullptr;
            int rc = graph_for(b, eps, &ex);
            if (rc) return rc;
            BCHK(b, hipGraphLaunch(ex, b->stream));
            b->cur ^= 1;
            t += eps;
            continue;
        }
        const bool fuse = fusable(b) && !last && n_rk_steps > 0;
        int rc = fuse ? QCBF_OK : enqueue_measure(b, b->dres, b->d_slot, true);
        if (rc) return rc;
        if (!last)
            for (int k = 0; k < n_rk_steps; k++) {
                rc = fuse && k == 0 ? enqueue_step(b, eps, b->cur, b->dres, b->d_slot, true)
                                    : enqueue_step(b, eps, b->cur);
                if (rc) return rc;
                b->cur ^= 1;
                t += eps;
            }
    }
    if (b->timing) BCHK(b, hipEventRecord(b->ev[1], b->stream));
    std::vector<double> r(need);
    BCHK(b, hipMemcpyAsync(r.data(), b->dres, need * sizeof(double), hipMemcpyDeviceToHost,
                           b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    if (b->timing) {
        float ms = 0.f;
        BCHK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        b->step_ms += ms;
        b->step_n += (long)(n_steps_gf - 1) * n_rk_steps;
    }
    // tQE[(c*n_steps_gf + n)*3 + {0,1,2}] = (t, Q, E); plaq[c*n_steps_gf + n]
    for (int n = 0; n < n_steps_gf; n++)
        for (int c = 0; c < b->n; c++) {
            const double* o = &r[3 * ((size_t)n * b->n + c)];
            double* q = tQE + 3 * ((size_t)c * n_steps_gf + n);
            q[0] = tvals[n];
            q[1] = o[2];
            q[2] = o[1];
            if (plaq) plaq[(size_t)c * n_steps_gf + n] = o[0];
        }
    return QCBF_OK;
}

extern "C" int qcbf_batch_exp_warnings(qcbf_batch* b, int* flagged) {
    CHECK_B(b);
    if (!flagged) return bfail(b, QCBF_ERR_ARG, "null output");
    int v = 0;
    BCHK(b, hipMemcpyAsync(&v, b->dwarn, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipMemsetAsync(b->dwarn, 0, sizeof(int), b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    *flagged = v;
    return QCBF_OK;
}

extern "C" int qcbf_batch_synchronize(qcbf_batch* b) {
    CHECK_B(b);
    BCHK(b, hipStreamSynchronize(b->stream));
    return QCBF_OK;
}

// ---- launch shape, graphs, timing -------------------------------------------------
extern "C" int qcbf_batch_tune(qcbf_batch* b, int key, long value) {
    CHECK_B(b);
    switch (key) {
        case QCBF_TUNE_MAP_MODE:
            if (value < 0 || value > 2) return bfail(b, QCBF_ERR_ARG, "map mode %ld not in 0..2", value);
            b->bl.map_mode = (int)value;
            break;
        case QCBF_TUNE_MAP_WIN:
            if (value < 8 || value % 8 != 0 || value > (1L << 24))
                return bfail(b, QCBF_ERR_ARG, "map window %ld not a multiple of 8 in [8, 2^24]", value);
            b->bl.map_win = (unsigned)value;
            break;
        case QCBF_TUNE_BLOCKS_PER_CU:
            if (value < 0 || value > 16) return bfail(b, QCBF_ERR_ARG, "blocks per CU %ld not in 0..16", value);
            b->blocks_per_cu = (int)value;
            break;
        case QCBF_TUNE_FUSED_MEASURE:
            if (value != 0 && value != 1) return bfail(b, QCBF_ERR_ARG, "fused measure %ld not 0 or 1", value);
            b->fuse_measure = value == 1;
            break;
        case QCBF_TUNE_THREADS_PER_SITE:
            if (value != 1 && value != 2 && value != 4)
                return bfail(b, QCBF_ERR_ARG, "threads per site %ld not 1, 2 or 4", value);
            b->bl.dpw = 4 / (int)value;
            break;
        case QCBF_TUNE_LDS_STAGE:
            if (value != 0 && value != 1) return bfail(b, QCBF_ERR_ARG, "LDS stage %ld not 0 or 1", value);
            b->lds_stage = (int)value;
            break;
        case QCBF_TUNE_TILE_SHAPE: {
            if (value != -1 && (value < 0 || value >= 4096))
                return bfail(b, QCBF_ERR_ARG, "tile shape %ld out of range", value);
            const int lg[4] = {(int)(value & 7), (int)((value >> 3) & 7), (int)((value >> 6) & 7),
                               (int)((value >> 9) & 7)};
            // kernels still queued on the handle's stream read the descriptors
            BCHK(b, hipStreamSynchronize(b->stream));
            for (auto& cd : b->cfgs) qcb::set_tile_shape(cd.g, value == -1 ? nullptr : lg);
            b->tile_shape = value;
            BCHK(b, hipMemcpy(b->d_cfgs, b->cfgs.data(), sizeof(cfg_desc) * b->n,
                              hipMemcpyHostToDevice));
            break;
        }
        default: return bfail(b, QCBF_ERR_ARG, "unknown tuning key %d", key);
    }
    set_grid(b);
    return QCBF_OK;
}

extern "C" int qcbf_batch_get_tune(const qcbf_batch* b, int key, long* value) {
    if (!b || !value) return QCBF_ERR_ARG;
    switch (key) {
        case QCBF_TUNE_MAP_MODE: *value = b->bl.map_mode; return QCBF_OK;
        case QCBF_TUNE_MAP_WIN: *value = b->bl.map_win; return QCBF_OK;
        case QCBF_TUNE_BLOCKS_PER_CU: *value = b->blocks_per_cu; return QCBF_OK;
        case QCBF_TUNE_TILE_SHAPE: *value = b->tile_shape; return QCBF_OK;
        case QCBF_TUNE_FUSED_MEASURE: *value = b->fuse_measure ? 1 : 0; return QCBF_OK;
        case QCBF_TUNE_THREADS_PER_SITE: *value = 4 / b->bl.dpw; return QCBF_OK;
        case QCBF_TUNE_LDS_STAGE: *value = b->lds_stage; return QCBF_OK;
        default: return QCBF_ERR_ARG;
    }
}

extern "C" int qcbf_batch_use_graph(qcbf_batch* b, int on) {
    CHECK_B(b);
    b->use_graph = on != 0;
    return QCBF_OK;
}

extern "C" int qcbf_batch_timing(qcbf_batch* b, int on, double* total_ms, long* rk3_steps) {
    CHECK_B(b);
    if (total_ms) *total_ms = b->step_ms;
    if (rk3_steps) *rk3_steps = b->step_n;
    if (on && !b->ev[0]) {
        BCHK(b, hipEventCreate(&b->ev[0]));
        BCHK(b, hipEventCreate(&b->ev[1]));
    }
    b->timing = on != 0;
    b->step_ms = 0.0;
    b->step_n = 0;
    return QCBF_OK;
}

// ---- autotuner ---------------------------------------------------------------------
// Candidate launch shapes are timed with HIP events over trial_steps RK3 steps
// of the real ensemble; the fields are saved first and restored afterwards, so
// tuning leaves the ensemble exactly as it was.
extern "C" int qcbf_batch_autotune(qcbf_batch* b, double eps, int trial_steps, long* mode,
                                   long* win, long* blocks_per_cu) {
    CHECK_B(b);
    if (trial_steps < 1) return bfail(b, QCBF_ERR_ARG, "trial_steps < 1");
    const size_t ub = b->link_words * wbytes(b), ab = b->alg_words * wbytes(b);
    char *saveU = nullptr, *saveA = nullptr;
    BCHK(b, hipMalloc(&saveU, ub));
    hipError_t e = hipMalloc(&saveA, ab);
    if (e != hipSuccess) {
        (void)hipFree(saveU);
        return bfail(b, QCBF_ERR_OOM, "autotune save buffer: %s", hipGetErrorString(e));
    }
    const int cur0 = b->cur;
    const std::vector<char> stale0 = b->rows_stale;
    // the trial steps may raise the exp warning: keep the caller's flag
    int warn0 = 0;
    e = hipMemcpyAsync(&warn0, b->dwarn, sizeof(int), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        (void)hipFree(saveU);
        (void)hipFree(saveA);
        return bfail(b, QCBF_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = QCBF_OK;
    struct cand {
        int mode;
        unsigned win;
        int bpc;
        long shape;
        int dpw;
    };
    // tile order x grid x tile shape at four threads per site, then two and
    // one thread per site on the best of those
    std::vector<cand> cands;
    const std::pair<int, unsigned> maps[] = {{0, 64}, {2, 256}, {2, 1024}};
    for (auto m : maps)
        for (int bpc : {0, 1, 2, 3, 4})
            for (long sh : {-1L, 594L, 83L}) cands.push_back({m.first, m.second, bpc, sh, 1});
    const size_t first_pass = cands.size();
    cand best = {b->bl.map_mode, b->bl.map_win, b->blocks_per_cu, b->tile_shape, b->bl.dpw};
    auto apply_shape = [&](long code) -> hipError_t {
        const int lg[4] = {(int)(code & 7), (int)((code >> 3) & 7), (int)((code >> 6) & 7),
                           (int)((code >> 9) & 7)};
        for (auto& cd : b->cfgs) qcb::set_tile_shape(cd.g, code == -1 ? nullptr : lg);
        b->tile_shape = code;
        return hipMemcpyAsync(b->d_cfgs, b->cfgs.data(), sizeof(cfg_desc) * b->n,
                              hipMemcpyHostToDevice, b->stream);
    };
    float best_ms = 1e30f;
    if ((e = hipMemcpyAsync(saveU, b->U[cur0], ub, hipMemcpyDeviceToDevice, b->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(saveA, b->acc, ab, hipMemcpyDeviceToDevice, b->stream)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        rc = bfail(b, QCBF_ERR_HIP, "autotune setup: %s", hipGetErrorString(e));
    for (size_t k = 0; rc == QCBF_OK && k < cands.size(); k++) {
        b->bl.map_mode = cands[k].mode;
        b->bl.map_win = cands[k].win;
        b->blocks_per_cu = cands[k].bpc;
        b->bl.dpw = cands[k].dpw;
        set_grid(b);
        if ((e = apply_shape(cands[k].shape)) != hipSuccess) {
            rc = bfail(b, QCBF_ERR_HIP, "autotune shape: %s", hipGetErrorString(e));
            break;
        }
        rc = enqueue_step(b, eps, b->cur);  // warm
        if (rc) break;
        b->cur ^= 1;
        if ((e = hipEventRecord(e0, b->stream)) != hipSuccess) {
            rc = bfail(b, QCBF_ERR_HIP, "event: %s", hipGetErrorString(e));
            break;
        }
        for (int s = 0; rc == QCBF_OK && s < trial_steps; s++) {
            rc = enqueue_step(b, eps, b->cur);
            b->cur ^= 1;
        }
        if (rc) break;
        float ms = 0.f;
        if ((e = hipEventRecord(e1, b->stream)) != hipSuccess ||
            (e = hipEventSynchronize(e1)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
            rc = bfail(b, QCBF_ERR_HIP, "autotune timing: %s", hipGetErrorString(e));
            break;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = cands[k];
        }
        if (k + 1 == first_pass)
            for (int d : {2, 4}) {
                cand c = best;
                c.dpw = d;
                cands.push_back(c);
            }
    }
    b->bl.map_mode = best.mode;
    b->bl.map_win = best.win;
    b->blocks_per_cu = best.bpc;
    b->bl.dpw = best.dpw;
    set_grid(b);
    if ((e = apply_shape(best.shape)) != hipSuccess && rc == QCBF_OK)
        rc = bfail(b, QCBF_ERR_HIP, "autotune shape: %s", hipGetErrorString(e));
    b->cur = cur0;
    b->rows_stale = stale0;
    if ((e = hipMemcpyAsync(b->U[cur0], saveU, ub, hipMemcpyDeviceToDevice, b->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b->acc, saveA, ab, hipMemcpyDeviceToDevice, b->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b->dwarn, &warn0, sizeof(int), hipMemcpyHostToDevice, b->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(b->stream)) != hipSuccess)
        if (rc == QCBF_OK) rc = bfail(b, QCBF_ERR_HIP, "autotune restore: %s", hipGetErrorString(e));
    (void)hipFree(saveU);
    (void)hipFree(saveA);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (mode) *mode = best.mode;
    if (win) *win = best.win;
    if (blocks_per_cu) *blocks_per_cu = best.bpc;
    return rc;
}
