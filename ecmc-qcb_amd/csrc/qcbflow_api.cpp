// qcbflow_api.cpp -- the C ABI of libqcbflow (include/qcbflow.h).
//
// Host orchestration of the flow: buffer ownership, RK3 sequencing
// (GradientFlow::rk3_step, src/flow/gradient_flow.cpp:124-135), the
// topo_charge_flowed cadence (src/observables/observables_mpi.cpp:159-186),
// ILDG I/O (src/io/ildg.cpp:34-237) and error mapping.  No CPU arithmetic
// path exists: every observable and every flowed link comes from the HIP
// kernels.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "ildg.hpp"
#include "kernels.hpp"
#include "qcbflow.h"

using qcb::lattice;

struct qcbf_event {
    int cls;
    hipEvent_t a, b;
};

struct qcbf_flow {
    int dev = 0;
    int precision = QCBF_FP64;
    qcb::prec prec = qcb::prec::f64;
    lattice g{};
    hipStream_t stream = nullptr;
    void* U[2] = {nullptr, nullptr};  // ping-pong link fields
    int cur = 0;                      // U[cur] is the flowing field
    bool rows_stale = false;          // U[cur] holds rows 0-1 only (a flow output)
    void* acc = nullptr;              // RK3 su(3) accumulator (8 reals / link)
    void* F[2] = {nullptr, nullptr};  // force_0, force_1 of the unfused API
    double* partials = nullptr;       // per-workgroup partial sums
    double* dres = nullptr;           // device results of measurements
    size_t dres_cap = 0;              // in doubles
    int* dwarn = nullptr;             // exp t > 1 flag
    int* dstuck = nullptr;            // heatbath retry-bound flag (kept apart from dwarn)
    unsigned hb_sweeps = 0;           // heatbath sweeps done (keys the random streams)
    void* stage = nullptr;            // device staging chunk for host <-> HBM transfers
    std::string err;
    // timing: every timing_every-th launch of each kernel class is bracketed
    bool timing = false;
    int timing_every = 1;
    long launch_no[QCBF_K_COUNT] = {};
    std::vector<qcbf_event> events;
    size_t ev_used = 0;
    double tot_ms[QCBF_K_COUNT] = {};
    long cnt[QCBF_K_COUNT] = {};
    bool fuse_measure = true;         // flow_measure: measurement inside the next stage 1
};

// Sites per host <-> HBM transfer chunk (37.7 MB of FP64 links).
constexpr unsigned kChunkSites = 65536;
constexpr size_t kChunkBytes = (size_t)kChunkSites * 72 * sizeof(double);

static thread_local std::string g_err;

static int fail(qcbf_flow* h, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_err = buf;
    return code;
}

int qcb::fail_without_handle(int code, const char* msg) { return fail(nullptr, code, "%s", msg); }

#define HIPCHK(h, expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (void)hipGetLastError(); /* the error is reported here, not by the next call */  \
            return fail((h), e_ == hipErrorOutOfMemory ? QCBF_ERR_OOM : QCBF_ERR_HIP,       \
                        "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
        }                                                                                   \
    } while (0)

#define CHECK_H(h)                                                    \
    do {                                                              \
        if (!(h)) return fail(nullptr, QCBF_ERR_ARG, "null handle"); \
        HIPCHK((h), hipSetDevice((h)->dev));                          \
    } while (0)

// ---- timing -----------------------------------------------------------------
static int timing_flush(qcbf_flow* h) {
    if (h->ev_used == 0) return QCBF_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < h->ev_used; k++) {
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, h->events[k].a, h->events[k].b));
        h->tot_ms[h->events[k].cls] += ms;
        h->cnt[h->events[k].cls] += 1;
    }
    h->ev_used = 0;
    return QCBF_OK;
}
static int timing_begin(qcbf_flow* h, int cls, qcbf_event** ev) {
    *ev = nullptr;
    if (!h->timing || h->launch_no[cls]++ % h->timing_every != 0) return QCBF_OK;
    if (h->ev_used == h->events.size()) {
        int rc = timing_flush(h);
        if (rc) return rc;
    }
    qcbf_event* e = &h->events[h->ev_used++];
    e->cls = cls;
    HIPCHK(h, hipEventRecord(e->a, h->stream));
    *ev = e;
    return QCBF_OK;
}
static int timing_end(qcbf_flow* h, qcbf_event* ev) {
    if (ev) HIPCHK(h, hipEventRecord(ev->b, h->stream));
    return QCBF_OK;
}

// ---- lifetime -----------------------------------------------------------------
static void release(qcbf_flow* h) {
    (void)hipSetDevice(h->dev);
    for (auto& e : h->events) {
        if (e.a) (void)hipEventDestroy(e.a);
        if (e.b) (void)hipEventDestroy(e.b);
    }
    (void)hipFree(h->U[0]);
    (void)hipFree(h->U[1]);
    (void)hipFree(h->acc);
    (void)hipFree(h->F[0]);
    (void)hipFree(h->F[1]);
    (void)hipFree(h->partials);
    (void)hipFree(h->dres);
    (void)hipFree(h->dwarn);
    (void)hipFree(h->dstuck);
    (void)hipFree(h->stage);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" const char* qcbf_last_error(const qcbf_flow* h) {
    return h ? h->err.c_str() : g_err.c_str();
}

static int ensure_results(qcbf_flow* h, size_t ndoubles) {
    if (ndoubles <= h->dres_cap) return QCBF_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    (void)hipFree(h->dres);
    h->dres = nullptr;
    h->dres_cap = 0;
    HIPCHK(h, hipMalloc(&h->dres, ndoubles * sizeof(double)));
    h->dres_cap = ndoubles;
    return QCBF_OK;
}

extern "C" int qcbf_create(const int dims[4], int precision, qcbf_flow** out) {
    if (!dims || !out) return fail(nullptr, QCBF_ERR_ARG, "null argument");
    *out = nullptr;
    if (precision != QCBF_FP64 && precision != QCBF_FP32_MIXED)
        return fail(nullptr, QCBF_ERR_ARG, "unknown precision %d", precision);
    unsigned long long V = 1;
    for (int d = 0; d < 4; d++) {
        if (dims[d] < 2 || dims[d] > 4096)
            return fail(nullptr, QCBF_ERR_ARG, "extent L[%d] = %d out of range [2, 4096]", d, dims[d]);
        V *= (unsigned long long)dims[d];
    }
    // below 2^28 sites: the FP32-mixed kernels address by 32-bit byte offsets
    static_assert((1ull << 28) <= qcb::kMaxSites32c, "volume limit beyond 32-bit site offsets");
    if (V >= (1ull << 28)) return fail(nullptr, QCBF_ERR_ARG, "volume %llu too large", V);
    qcbf_flow* h = new qcbf_flow();
    h->precision = precision;
    h->prec = precision == QCBF_FP64 ? qcb::prec::f64 : qcb::prec::f32m;
    h->g = qcb::make_lattice(dims);
    hipError_t e = hipGetDevice(&h->dev);
    if (e != hipSuccess) {
        delete h;
        return fail(nullptr, QCBF_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    }
#define CREATE_CHK(expr)                                                                      \
    do {                                                                                      \
        hipError_t e2 = (expr);                                                               \
        if (e2 != hipSuccess) {                                                               \
            (void)hipGetLastError();                                                          \
            release(h);                                                                       \
            return fail(nullptr, e2 == hipErrorOutOfMemory ? QCBF_ERR_OOM : QCBF_ERR_HIP,     \
                        "%s: %s", #expr, hipGetErrorString(e2));                              \
        }                                                                                     \
    } while (0)
    CREATE_CHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    CREATE_CHK(hipMalloc(&h->U[0], qcb::link_field_bytes(h->prec, h->g.V)));
    CREATE_CHK(hipMalloc(&h->U[1], qcb::link_field_bytes(h->prec, h->g.V)));
    CREATE_CHK(hipMalloc(&h->acc, qcb::alg_field_bytes(h->prec, h->g.V)));
    CREATE_CHK(hipMalloc(&h->partials, (size_t)qcb::measure_blocks(h->g) * 3 * sizeof(double)));
    CREATE_CHK(hipMalloc(&h->dwarn, sizeof(int)));
    CREATE_CHK(hipMemsetAsync(h->dwarn, 0, sizeof(int), h->stream));
    CREATE_CHK(hipMalloc(&h->dstuck, sizeof(int)));
    CREATE_CHK(hipMalloc(&h->dres, 64 * sizeof(double)));
    h->dres_cap = 64;
    if (qcb::launch_cold_start(h->prec, h->U[0], h->g, h->stream) != hipSuccess) {
        release(h);
        return fail(nullptr, QCBF_ERR_HIP, "cold start launch failed");
    }
    CREATE_CHK(hipStreamSynchronize(h->stream));
#undef CREATE_CHK
    *out = h;
    return QCBF_OK;
}

extern "C" int qcbf_destroy(qcbf_flow* h) {
    if (!h) return QCBF_OK;
    (void)hipSetDevice(h->dev);
    (void)hipStreamSynchronize(h->stream);
    release(h);
    return QCBF_OK;
}

extern "C" int qcbf_get_dims(const qcbf_flow* h, int dims[4], int* precision) {
    if (!h || !dims) return fail(nullptr, QCBF_ERR_ARG, "null argument");
    for (int d = 0; d < 4; d++) dims[d] = h->g.L[d];
    if (precision) *precision = h->precision;
    return QCBF_OK;
}

// ---- links in / out ------------------------------------------------------------
static int ensure_stage(qcbf_flow* h) {
    if (!h->stage) HIPCHK(h, hipMalloc(&h->stage, kChunkBytes));
    return QCBF_OK;
}

// Complete the third rows of a flow output before the field is read whole.
static int complete_rows(qcbf_flow* h) {
    if (!h->rows_stale) return QCBF_OK;
    HIPCHK(h, qcb::launch_complete_rows(h->prec, h->U[h->cur], h->g.V, h->stream));
    h->rows_stale = false;
    return QCBF_OK;
}

extern "C" int qcbf_set_links(qcbf_flow* h, const double* links) {
    CHECK_H(h);
    if (!links) return fail(h, QCBF_ERR_ARG, "null links");
    int rc = ensure_stage(h);
    if (rc) return rc;
    for (unsigned s0 = 0; s0 < h->g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, h->g.V - s0);
        HIPCHK(h, hipMemcpyAsync(h->stage, links + (size_t)s0 * 72, (size_t)n * 72 * sizeof(double),
                                 hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, qcb::launch_aos_to_field(h->prec, h->stage, 0, h->U[h->cur], h->g, s0, n,
                                           h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->rows_stale = false;
    return QCBF_OK;
}

extern "C" int qcbf_get_links(qcbf_flow* h, double* links) {
    CHECK_H(h);
    if (!links) return fail(h, QCBF_ERR_ARG, "null links");
    int rc = ensure_stage(h);
    if (rc || (rc = complete_rows(h))) return rc;
    double* st = static_cast<double*>(h->stage);
    for (unsigned s0 = 0; s0 < h->g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, h->g.V - s0);
        HIPCHK(h, qcb::launch_field_to_aos(h->prec, h->U[h->cur], st, h->g, s0, n, false,
                                           h->stream));
        HIPCHK(h, hipMemcpyAsync(links + (size_t)s0 * 72, st, (size_t)n * 72 * sizeof(double),
                                 hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QCBF_OK;
}

extern "C" int qcbf_copy_field(qcbf_flow* dst, qcbf_flow* src) {
    CHECK_H(dst);
    if (!src) return fail(dst, QCBF_ERR_ARG, "null source handle");
    if (src->dev != dst->dev) return fail(dst, QCBF_ERR_ARG, "handles on different devices");
    for (int d = 0; d < 4; d++)
        if (src->g.L[d] != dst->g.L[d]) return fail(dst, QCBF_ERR_ARG, "extents differ");
    if (src == dst) return QCBF_OK;
    int rc = ensure_stage(dst);
    if (rc) return rc;
    if ((rc = complete_rows(src))) return fail(dst, rc, "source: %s", src->err.c_str());
    // the source's queued work must be done before its field is read
    HIPCHK(dst, hipStreamSynchronize(src->stream));
    double* st = static_cast<double*>(dst->stage);
    for (unsigned s0 = 0; s0 < dst->g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, dst->g.V - s0);
        HIPCHK(dst, qcb::launch_field_to_aos(src->prec, src->U[src->cur], st, src->g, s0, n, false,
                                             dst->stream));
        HIPCHK(dst, qcb::launch_aos_to_field(dst->prec, st, 0, dst->U[dst->cur], dst->g, s0, n,
                                             dst->stream));
    }
    HIPCHK(dst, hipStreamSynchronize(dst->stream));
    dst->rows_stale = false;
    return QCBF_OK;
}

// The ensemble ABI copies from a single-lattice handle through this view.
int qcb::field_view(qcbf_flow* h, int* prec, const void** field, hipStream_t* stream) {
    if (!h) return QCBF_ERR_ARG;
    const int rc = complete_rows(h);  // the field is read whole
    if (rc) return rc;
    *prec = (int)h->prec;
    *field = h->U[h->cur];
    *stream = h->stream;
    return QCBF_OK;
}

extern "C" int qcbf_hot_start(qcbf_flow* h, uint64_t seed) {
    CHECK_H(h);
    HIPCHK(h, qcb::launch_hot_start(h->prec, h->U[h->cur], h->g, seed, h->stream));
    h->rows_stale = false;
    return QCBF_OK;
}

extern "C" int qcbf_shift(qcbf_flow* h, const int shift[4]) {
    CHECK_H(h);
    if (!shift) return fail(h, QCBF_ERR_ARG, "null shift");
    int sh[4];
    for (int d = 0; d < 4; d++) sh[d] = ((shift[d] % h->g.L[d]) + h->g.L[d]) % h->g.L[d];
    HIPCHK(h, qcb::launch_shift(h->prec, h->U[h->cur], h->U[h->cur ^ 1], h->g, sh, h->stream));
    h->cur ^= 1;  // the stored words move as they are; rows_stale carries over
    return QCBF_OK;
}

extern "C" int qcbf_heatbath(qcbf_flow* h, double beta, int n_sweeps, uint64_t seed) {
    CHECK_H(h);
    if (h->prec != qcb::prec::f64)
        return fail(h, QCBF_ERR_STATE, "the heatbath runs on FP64 fields only");
    for (int d = 0; d < 4; d++)
        if (h->g.L[d] % 2)
            return fail(h, QCBF_ERR_ARG, "the checkerboard heatbath needs even extents (L[%d] = %d)",
                        d, h->g.L[d]);
    if (n_sweeps < 0 || !(beta > 0.0)) return fail(h, QCBF_ERR_ARG, "bad heatbath arguments");
    HIPCHK(h, hipMemsetAsync(h->dstuck, 0, sizeof(int), h->stream));
    HIPCHK(h, qcb::launch_heatbath_f64(static_cast<double2*>(h->U[h->cur]), h->g, beta, n_sweeps,
                                       seed, h->hb_sweeps, h->dstuck, h->stream));
    h->hb_sweeps += (unsigned)n_sweeps;
    if (n_sweeps > 0) h->rows_stale = false;  // every link rewritten whole
    int stuck = 0;
    HIPCHK(h, hipMemcpyAsync(&stuck, h->dstuck, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (stuck) return fail(h, QCBF_ERR_STATE, "an SU(2) heatbath draw hit its retry bound");
    return QCBF_OK;
}

extern "C" int qcbf_cold_start(qcbf_flow* h) {
    CHECK_H(h);
    HIPCHK(h, qcb::launch_cold_start(h->prec, h->U[h->cur], h->g, h->stream));
    h->rows_stale = false;
    return QCBF_OK;
}

// ---- ILDG ----------------------------------------------------------------------
namespace {
struct pinned_pair {
    void* p[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~pinned_pair() {
        for (int k = 0; k < 2; k++) {
            if (ev[k]) {
                (void)hipEventSynchronize(ev[k]);
                (void)hipEventDestroy(ev[k]);
            }
            if (p[k]) (void)hipHostFree(p[k]);
        }
    }
};
struct file_closer {
    FILE* fp;
    ~file_closer() {
        if (fp) fclose(fp);
    }
};
}  // namespace

extern "C" int qcbf_load_ildg(qcbf_flow* h, const char* path) {
    CHECK_H(h);
    if (!path) return fail(h, QCBF_ERR_ARG, "null path");
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(h, QCBF_ERR_IO, "cannot open '%s' for reading", path);
    file_closer fc{fp};
    qcb::ildg_info info;
    std::string msg;
    int rc = qcb::ildg_scan(fp, &info, &msg);
    if (rc) return fail(h, rc, "%s: %s", path, msg.c_str());
    if (info.has_format) {
        for (int d = 0; d < 4; d++)
            if (info.dims[d] != h->g.L[d])
                return fail(h, QCBF_ERR_FORMAT,
                            "%s: ildg-format extents %dx%dx%dx%d differ from the handle's %dx%dx%dx%d",
                            path, info.dims[0], info.dims[1], info.dims[2], info.dims[3], h->g.L[0],
                            h->g.L[1], h->g.L[2], h->g.L[3]);
        if (info.precision != 64 && info.precision != 32)
            return fail(h, QCBF_ERR_FORMAT, "%s: unsupported precision %d", path, info.precision);
    }
    const size_t word = info.precision == 32 ? 4 : 8;
    const unsigned long long expect = (unsigned long long)h->g.V * 72 * word;
    if (info.data_bytes != expect)
        return fail(h, QCBF_ERR_FORMAT, "%s: binary size %llu bytes, expected %llu", path,
                    info.data_bytes, expect);
    // Stream the payload through two pinned buffers and a device staging chunk;
    // each chunk is byte-swapped and transposed into the field on the device.
    int rc2 = ensure_stage(h);
    if (rc2) return rc2;
    const size_t site_bytes = 72 * word;
    const unsigned chunk_sites = (unsigned)(kChunkBytes / site_bytes);
    pinned_pair pp;
    for (int k = 0; k < 2; k++) {
        HIPCHK(h, hipHostMalloc(&pp.p[k], kChunkBytes, hipHostMallocDefault));
        HIPCHK(h, hipEventCreateWithFlags(&pp.ev[k], hipEventDisableTiming));
    }
    if (fseeko(fp, info.data_offset, SEEK_SET) != 0)
        return fail(h, QCBF_ERR_IO, "%s: seek failed", path);
    int k = 0;
    bool used[2] = {false, false};
    for (unsigned s0 = 0; s0 < h->g.V; s0 += chunk_sites) {
        const unsigned n = std::min(chunk_sites, h->g.V - s0);
        const size_t nb = (size_t)n * site_bytes;
        if (used[k]) HIPCHK(h, hipEventSynchronize(pp.ev[k]));
        if (fread(pp.p[k], 1, nb, fp) != nb) return fail(h, QCBF_ERR_IO, "%s: short read", path);
        HIPCHK(h, hipMemcpyAsync(h->stage, pp.p[k], nb, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, qcb::launch_aos_to_field(h->prec, h->stage, word == 8 ? 1 : 2, h->U[h->cur], h->g,
                                           s0, n, h->stream));
        HIPCHK(h, hipEventRecord(pp.ev[k], h->stream));
        used[k] = true;
        k ^= 1;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->rows_stale = false;
    return QCBF_OK;
}

extern "C" int qcbf_save_ildg(qcbf_flow* h, const char* path) {
    CHECK_H(h);
    if (!path) return fail(h, QCBF_ERR_ARG, "null path");
    int rc = ensure_stage(h);
    if (rc || (rc = complete_rows(h))) return rc;
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(h, QCBF_ERR_IO, "cannot open '%s' for writing", path);
    file_closer fc{fp};
    const std::string xml = qcb::ildg_format_xml(h->g.L, 64);
    const uint64_t nbin = (uint64_t)h->g.V * 72 * sizeof(double);
    bool ok = qcb::lime_write_header(fp, "ildg-format", xml.size(), true, true);
    ok = ok && fwrite(xml.data(), 1, xml.size(), fp) == xml.size();
    static const char zeros[8] = {0};
    const size_t pad = (8 - xml.size() % 8) % 8;
    ok = ok && fwrite(zeros, 1, pad, fp) == pad;
    ok = ok && qcb::lime_write_header(fp, "ildg-binary-data", nbin, false, true);
    if (!ok) return fail(h, QCBF_ERR_IO, "%s: short write of LIME headers", path);
    pinned_pair pp;
    HIPCHK(h, hipHostMalloc(&pp.p[0], kChunkBytes, hipHostMallocDefault));
    double* st = static_cast<double*>(h->stage);
    for (unsigned s0 = 0; s0 < h->g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, h->g.V - s0);
        const size_t nb = (size_t)n * 72 * sizeof(double);
        HIPCHK(h, qcb::launch_field_to_aos(h->prec, h->U[h->cur], st, h->g, s0, n, true, h->stream));
        HIPCHK(h, hipMemcpyAsync(pp.p[0], st, nb, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (fwrite(pp.p[0], 1, nb, fp) != nb) return fail(h, QCBF_ERR_IO, "%s: short write", path);
    }
    if (fflush(fp) != 0) return fail(h, QCBF_ERR_IO, "%s: flush failed", path);
    return QCBF_OK;
}

// ---- flow ------------------------------------------------------------------------
// One RK3 step.  With measure_out set, stage 1 also measures its input field
// into measure_out[0..2] (the fused kernel).
static int rk3_once(qcbf_flow* h, double eps, double* measure_out = nullptr) {
    const double cz[3] = {0.25 * eps, (8.0 / 9.0) * eps, 0.75 * eps};
    const double ca[3] = {0.0, (17.0 / 36.0) * eps, eps};
    for (int stage = 1; stage <= 3; stage++) {
        qcbf_event* ev;
        const bool fused = stage == 1 && measure_out;
        int rc = timing_begin(h, fused ? QCBF_K_STAGE1_MEASURE : QCBF_K_STAGE1 + stage - 1, &ev);
        if (rc) return rc;
        if (fused)
            HIPCHK(h, qcb::launch_rk3_stage1_measure(h->prec, h->U[h->cur], h->U[h->cur ^ 1],
                                                     h->acc, h->g, cz[0], h->dwarn, h->partials,
                                                     measure_out, h->stream));
        else if (h->g.lds_stage && qcb::stage_lds_applies(h->prec, h->g))
            HIPCHK(h, qcb::launch_rk3_stage_lds(h->prec, stage, h->U[h->cur], h->U[h->cur ^ 1], h->acc,
                                                h->g, cz[stage - 1], ca[stage - 1], h->dwarn,
                                                h->stream));
        else
            HIPCHK(h, qcb::launch_rk3_stage(h->prec, stage, h->U[h->cur], h->U[h->cur ^ 1],
                                            h->acc, h->g, cz[stage - 1], ca[stage - 1], h->dwarn,
                                            h->stream));
        rc = timing_end(h, ev);
        if (rc) return rc;
        h->cur ^= 1;
    }
    h->rows_stale = h->prec == qcb::prec::f64 && qcb::kF64FlowStoresTwoRows;
    return QCBF_OK;
}

extern "C" int qcbf_rk3_step(qcbf_flow* h, double eps, int n_steps) {
    CHECK_H(h);
    if (n_steps < 0) return fail(h, QCBF_ERR_ARG, "n_steps < 0");
    for (int k = 0; k < n_steps; k++) {
        int rc = rk3_once(h, eps);
        if (rc) return rc;
    }
    return QCBF_OK;
}

static int ensure_force(qcbf_flow* h, int i) {
    if (!h->F[i]) HIPCHK(h, hipMalloc(&h->F[i], qcb::alg_field_bytes(h->prec, h->g.V)));
    return QCBF_OK;
}

extern "C" int qcbf_compute_force(qcbf_flow* h, int i) {
    CHECK_H(h);
    if (i != 0 && i != 1) return fail(h, QCBF_ERR_ARG, "force index %d not 0 or 1", i);
    int rc = ensure_force(h, i);
    if (rc) return rc;
    HIPCHK(h, qcb::launch_force(h->prec, h->U[h->cur], h->F[i], h->g, h->stream));
    return QCBF_OK;
}

static int update(qcbf_flow* h, int which, double eps) {
    if (!h->F[0] || (which > 1 && !h->F[1]))
        return fail(h, QCBF_ERR_STATE, "compute_w%d needs the forces it combines", which);
    HIPCHK(h, qcb::launch_update(h->prec, which, h->U[h->cur], h->F[0], h->F[1], h->g, eps,
                                 h->dwarn, h->stream));
    h->rows_stale = h->prec == qcb::prec::f64 && qcb::kF64FlowStoresTwoRows;
    return QCBF_OK;
}
extern "C" int qcbf_compute_w1(qcbf_flow* h, double eps) {
    CHECK_H(h);
    return update(h, 1, eps);
}
extern "C" int qcbf_compute_w2(qcbf_flow* h, double eps) {
    CHECK_H(h);
    return update(h, 2, eps);
}
extern "C" int qcbf_compute_w3(qcbf_flow* h, double eps) {
    CHECK_H(h);
    return update(h, 3, eps);
}
extern "C" int qcbf_replace_force_0(qcbf_flow* h) {
    CHECK_H(h);
    if (!h->F[0] || !h->F[1]) return fail(h, QCBF_ERR_STATE, "replace_force_0 needs both forces");
    HIPCHK(h, qcb::launch_replace_force0(h->prec, h->F[0], h->F[1], h->g, h->stream));
    return QCBF_OK;
}
extern "C" int qcbf_get_force(qcbf_flow* h, int i, double* out) {
    CHECK_H(h);
    if ((i != 0 && i != 1) || !out) return fail(h, QCBF_ERR_ARG, "bad force index or null output");
    if (!h->F[i]) return fail(h, QCBF_ERR_STATE, "force_%d has not been computed", i);
    int rc = ensure_stage(h);
    if (rc) return rc;
    // through the staging chunk, kChunkSites sites at a time, as qcbf_get_links
    double* st = static_cast<double*>(h->stage);
    for (unsigned s0 = 0; s0 < h->g.V; s0 += kChunkSites) {
        const unsigned n = std::min(kChunkSites, h->g.V - s0);
        HIPCHK(h, qcb::launch_alg_to_aos(h->prec, h->F[i], st, h->g, s0, n, h->stream));
        HIPCHK(h, hipMemcpyAsync(out + (size_t)s0 * 72, st, (size_t)n * 72 * sizeof(double),
                                 hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QCBF_OK;
}

// ---- observables ---------------------------------------------------------------
static int measure_into(qcbf_flow* h, double* dout, bool clover) {
    qcbf_event* ev;
    int rc = timing_begin(h, QCBF_K_MEASURE, &ev);
    if (rc) return rc;
    HIPCHK(h, qcb::launch_measure(h->prec, h->U[h->cur], h->g, h->partials, dout, clover, h->stream));
    return timing_end(h, ev);
}

static int measure_now(qcbf_flow* h, bool clover, double out[3]) {
    int rc = measure_into(h, h->dres, clover);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(out, h->dres, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QCBF_OK;
}

extern "C" int qcbf_plaquette(qcbf_flow* h, double* plaq) {
    CHECK_H(h);
    if (!plaq) return fail(h, QCBF_ERR_ARG, "null output");
    double r[3];
    int rc = measure_now(h, false, r);
    if (rc) return rc;
    *plaq = r[0];
    return QCBF_OK;
}

extern "C" int qcbf_q_e_clover(qcbf_flow* h, double* Q, double* E) {
    CHECK_H(h);
    if (!Q || !E) return fail(h, QCBF_ERR_ARG, "null output");
    double r[3];
    int rc = measure_now(h, true, r);
    if (rc) return rc;
    *E = r[1];
    *Q = r[2];
    return QCBF_OK;
}

extern "C" int qcbf_measure(qcbf_flow* h, double* plaq, double* E, double* Q) {
    CHECK_H(h);
    if (!plaq || !Q || !E) return fail(h, QCBF_ERR_ARG, "null output");
    double r[3];
    int rc = measure_now(h, true, r);
    if (rc) return rc;
    *plaq = r[0];
    *E = r[1];
    *Q = r[2];
    return QCBF_OK;
}

extern "C" int qcbf_clover_site(qcbf_flow* h, long site, int mu, int nu, double* out) {
    CHECK_H(h);
    if (!out || site < 0 || (unsigned long)site >= h->g.V || mu < 0 || mu > 3 || nu < 0 || nu > 3)
        return fail(h, QCBF_ERR_ARG, "clover_site: bad site, direction or null output");
    HIPCHK(h, qcb::launch_clover_site(h->prec, h->U[h->cur], h->g, (unsigned)site, mu, nu, h->dres,
                                      h->stream));
    HIPCHK(h, hipMemcpyAsync(out, h->dres, 18 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QCBF_OK;
}

extern "C" int qcbf_local_q_e(qcbf_flow* h, long site, double* q, double* e) {
    CHECK_H(h);
    if (!q || !e || site < 0 || (unsigned long)site >= h->g.V)
        return fail(h, QCBF_ERR_ARG, "local_q_e: bad site or null output");
    HIPCHK(h, qcb::launch_local_q_e(h->prec, h->U[h->cur], h->g, (unsigned)site, h->dres, h->stream));
    double r[2];
    HIPCHK(h, hipMemcpyAsync(r, h->dres, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *q = r[0];
    *e = r[1];
    return QCBF_OK;
}

extern "C" int qcbf_flow_measure(qcbf_flow* h, double eps, int n_steps_gf, int n_rk_steps,
                                 double* tQE, double* plaq) {
    CHECK_H(h);
    if (n_steps_gf < 0 || n_rk_steps < 0 || (n_steps_gf > 0 && !tQE))
        return fail(h, QCBF_ERR_ARG, "bad flow_measure arguments");
    if (n_steps_gf == 0) return QCBF_OK;
    int rc = ensure_results(h, (size_t)n_steps_gf * 3);
    if (rc) return rc;
    std::vector<double> tvals(n_steps_gf);
    double t = 0.0;
    // Every measurement but the last is followed by an RK3 step on the same
    // field: its stage 1 takes the measurement along (same results).
    const bool fuse =
        h->fuse_measure && qcb::kFusedMeasureBuilt && h->g.lg_tile[0] < 0 && n_rk_steps > 0;
    for (int n = 0; n < n_steps_gf; n++) {
        tvals[n] = t;
        double* out = h->dres + 3 * (size_t)n;
        const bool last = n == n_steps_gf - 1;
        if (!(fuse && !last)) {
            rc = measure_into(h, out, true);
            if (rc) return rc;
        }
        if (!last) {
            for (int k = 0; k < n_rk_steps; k++) {
                rc = rk3_once(h, eps, fuse && k == 0 ? out : nullptr);
                if (rc) return rc;
                t += eps;
            }
        }
    }
    std::vector<double> r((size_t)n_steps_gf * 3);
    HIPCHK(h, hipMemcpyAsync(r.data(), h->dres, r.size() * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int n = 0; n < n_steps_gf; n++) {
        tQE[3 * n] = tvals[n];
        tQE[3 * n + 1] = r[3 * n + 2];
        tQE[3 * n + 2] = r[3 * n + 1];
        if (plaq) plaq[n] = r[3 * n];
    }
    return QCBF_OK;
}

extern "C" int qcbf_exp_warnings(qcbf_flow* h, int* flagged) {
    CHECK_H(h);
    if (!flagged) return fail(h, QCBF_ERR_ARG, "null output");
    int v = 0;
    HIPCHK(h, hipMemcpyAsync(&v, h->dwarn, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemsetAsync(h->dwarn, 0, sizeof(int), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *flagged = v;
    return QCBF_OK;
}

extern "C" int qcbf_unitarity(qcbf_flow* h, double* max_dev, double* max_det_dev) {
    CHECK_H(h);
    if (!max_dev || !max_det_dev) return fail(h, QCBF_ERR_ARG, "null output");
    HIPCHK(h, qcb::launch_unitarity(h->prec, h->U[h->cur], h->g, h->partials, h->dres, h->stream));
    double r[2];
    HIPCHK(h, hipMemcpyAsync(r, h->dres, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *max_dev = r[0];
    *max_det_dev = r[1];
    return QCBF_OK;
}

extern "C" int qcbf_synchronize(qcbf_flow* h) {
    CHECK_H(h);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QCBF_OK;
}

// ---- exp on independent matrices -------------------------------------------------
extern "C" int qcbf_su3_exp(const double* Z, int n, double coeff, double* out, int* flagged) {
    if (n < 0 || (n > 0 && (!Z || !out))) return fail(nullptr, QCBF_ERR_ARG, "bad arguments");
    if (flagged) *flagged = 0;
    if (n == 0) return QCBF_OK;
    const size_t bytes = (size_t)n * 18 * sizeof(double);
    double *dz = nullptr, *dout = nullptr;
    int* dw = nullptr;
    int rc = QCBF_OK;
    hipError_t e;
    auto chk = [&](hipError_t x, const char* what) {
        if (x != hipSuccess && rc == QCBF_OK)
            rc = fail(nullptr, x == hipErrorOutOfMemory ? QCBF_ERR_OOM : QCBF_ERR_HIP, "%s: %s",
                      what, hipGetErrorString(x));
        return x == hipSuccess;
    };
    if (chk(hipMalloc(&dz, bytes), "hipMalloc") && chk(hipMalloc(&dout, bytes), "hipMalloc") &&
        chk(hipMalloc(&dw, sizeof(int)), "hipMalloc") && chk(hipMemset(dw, 0, sizeof(int)), "hipMemset") &&
        chk(hipMemcpy(dz, Z, bytes, hipMemcpyHostToDevice), "hipMemcpy") &&
        chk(qcb::launch_exp_batch_f64(dz, dout, n, coeff, dw, nullptr), "exp kernel") &&
        chk(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost), "hipMemcpy")) {
        int w = 0;
        e = hipMemcpy(&w, dw, sizeof(int), hipMemcpyDeviceToHost);
        if (chk(e, "hipMemcpy") && flagged) *flagged = w;
    }
    (void)hipFree(dz);
    (void)hipFree(dout);
    (void)hipFree(dw);
    return rc;
}

// ---- launch tuning -------------------------------------------------------------------
extern "C" int qcbf_tune(qcbf_flow* h, int key, long value) {
    CHECK_H(h);
    switch (key) {
        case QCBF_TUNE_MAP_MODE:
            if (value < 0 || value > 3) return fail(h, QCBF_ERR_ARG, "map mode %ld not in 0..3", value);
            h->g.map_mode = (int)value;
            return QCBF_OK;
        case QCBF_TUNE_MAP_WIN:
            if (value < 1 || value > (1L << 24))
                return fail(h, QCBF_ERR_ARG, "map window %ld not in [1, 2^24]", value);
            h->g.map_win = (unsigned)value;
            return QCBF_OK;
        case QCBF_TUNE_FUSED_MEASURE:
            if (value != 0 && value != 1) return fail(h, QCBF_ERR_ARG, "fused measure %ld not 0 or 1", value);
            h->fuse_measure = value == 1;
            return QCBF_OK;
        case QCBF_TUNE_THREADS_PER_SITE:
            if (value != 1 && value != 2 && value != 4)
                return fail(h, QCBF_ERR_ARG, "threads per site %ld not 1, 2 or 4", value);
            h->g.dpw = 4 / (int)value;
            return QCBF_OK;
        case QCBF_TUNE_LDS_STAGE:
            if (value != 0 && value != 1) return fail(h, QCBF_ERR_ARG, "LDS stage %ld not 0 or 1", value);
            h->g.lds_stage = (int)value;
            return QCBF_OK;
        case QCBF_TUNE_TILE_SHAPE: {
            if (value == -1) {
                qcb::set_tile_shape(h->g, nullptr);
                return QCBF_OK;
            }
            const int lg[4] = {(int)(value & 7), (int)((value >> 3) & 7), (int)((value >> 6) & 7),
                               (int)((value >> 9) & 7)};
            if (value < 0 || value >= 4096 || !qcb::set_tile_shape(h->g, lg)) {
                qcb::set_tile_shape(h->g, nullptr);
                return fail(h, QCBF_ERR_ARG, "tile shape %ld does not tile the lattice", value);
            }
            return QCBF_OK;
        }
        default:
            return fail(h, QCBF_ERR_ARG, "unknown tuning key %d", key);
    }
}

static long tile_code(const qcb::lattice& g) {
    if (g.lg_tile[0] < 0) return -1;
    return g.lg_tile[0] + 8L * g.lg_tile[1] + 64L * g.lg_tile[2] + 512L * g.lg_tile[3];
}

extern "C" int qcbf_get_tune(const qcbf_flow* h, int key, long* value) {
    if (!h || !value) return fail(nullptr, QCBF_ERR_ARG, "null argument");
    switch (key) {
        case QCBF_TUNE_MAP_MODE: *value = h->g.map_mode; return QCBF_OK;
        case QCBF_TUNE_MAP_WIN: *value = h->g.map_win; return QCBF_OK;
        case QCBF_TUNE_TILE_SHAPE: *value = tile_code(h->g); return QCBF_OK;
        case QCBF_TUNE_FUSED_MEASURE: *value = h->fuse_measure ? 1 : 0; return QCBF_OK;
        case QCBF_TUNE_THREADS_PER_SITE: *value = 4 / h->g.dpw; return QCBF_OK;
        case QCBF_TUNE_LDS_STAGE: *value = h->g.lds_stage; return QCBF_OK;
        default: return fail(nullptr, QCBF_ERR_ARG, "unknown tuning key %d", key);
    }
}

// ---- measurement support ------------------------------------------------------------
extern "C" int qcbf_timing_enable(qcbf_flow* h, int on) {
    CHECK_H(h);
    if (on < 0) return fail(h, QCBF_ERR_ARG, "timing_enable: %d < 0", on);
    int rc = timing_flush(h);
    if (rc) return rc;
    if (on && h->events.empty()) {
        h->events.resize(4096);
        for (auto& ev : h->events) ev.a = ev.b = nullptr;
        for (auto& ev : h->events) {
            HIPCHK(h, hipEventCreate(&ev.a));
            HIPCHK(h, hipEventCreate(&ev.b));
        }
    }
    h->timing = on != 0;
    h->timing_every = on > 0 ? on : 1;
    for (int k = 0; k < QCBF_K_COUNT; k++) {
        h->tot_ms[k] = 0.0;
        h->cnt[k] = 0;
        h->launch_no[k] = 0;
    }
    return QCBF_OK;
}

extern "C" int qcbf_timing_read(qcbf_flow* h, double total_ms[QCBF_K_COUNT],
                                long launches[QCBF_K_COUNT]) {
    CHECK_H(h);
    int rc = timing_flush(h);
    if (rc) return rc;
    for (int k = 0; k < QCBF_K_COUNT; k++) {
        if (total_ms) total_ms[k] = h->tot_ms[k];
        if (launches) launches[k] = h->cnt[k];
    }
    return QCBF_OK;
}

extern "C" int qcbf_copy_bench(qcbf_flow* h, size_t bytes, int reps, double* ms) {
    CHECK_H(h);
    if (!ms || reps <= 0 || bytes < 16) return fail(h, QCBF_ERR_ARG, "bad copy_bench arguments");
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    HIPCHK(h, hipMalloc(&a, n * sizeof(double2)));
    hipError_t e = hipMalloc(&b, n * sizeof(double2));
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(h, QCBF_ERR_OOM, "hipMalloc: %s", hipGetErrorString(e));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float t = 0.f;
    // the first failing call names the error; later steps are skipped
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * sizeof(double2), h->stream);
    if (e == hipSuccess) e = qcb::launch_copy(a, b, n, h->stream);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    for (int r = 0; e == hipSuccess && r < reps; r++) e = qcb::launch_copy(a, b, n, h->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e != hipSuccess) (void)hipStreamSynchronize(h->stream);  // nothing in flight on a, b
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(h, QCBF_ERR_HIP, "copy bench: %s", hipGetErrorString(e));
    *ms = (double)t / reps;
    return QCBF_OK;
}

extern "C" int qcbf_device_info(char* name, int name_len, int* compute_units) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, QCBF_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    hipDeviceProp_t p;
    e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess)
        return fail(nullptr, QCBF_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (name && name_len > 0) snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    if (compute_units) *compute_units = p.multiProcessorCount;
    return QCBF_OK;
}

// ---- autotuner ---------------------------------------------------------------------
// Each candidate tile order is timed with HIP events over trial_steps RK3 steps
// on the real field, which is saved first and restored afterwards.
extern "C" int qcbf_autotune(qcbf_flow* h, double eps, int trial_steps, long* mode, long* win) {
    CHECK_H(h);
    if (trial_steps < 1) return fail(h, QCBF_ERR_ARG, "trial_steps < 1");
    const size_t ub = qcb::link_field_bytes(h->prec, h->g.V);
    const size_t ab = qcb::alg_field_bytes(h->prec, h->g.V);
    void *saveU = nullptr, *saveA = nullptr;
    HIPCHK(h, hipMalloc(&saveU, ub));
    hipError_t e = hipMalloc(&saveA, ab);
    if (e != hipSuccess) {
        (void)hipFree(saveU);
        return fail(h, QCBF_ERR_OOM, "autotune save buffer: %s", hipGetErrorString(e));
    }
    const int cur0 = h->cur;
    const bool timing0 = h->timing, stale0 = h->rows_stale;
    h->timing = false;
    // the trial steps may raise the exp warning: keep the caller's flag
    int warn0 = 0;
    e = hipMemcpyAsync(&warn0, h->dwarn, sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        (void)hipFree(saveU);
        (void)hipFree(saveA);
        h->timing = timing0;
        return fail(h, QCBF_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    }
    struct cand {
        int mode;
        unsigned win;
        int lg[4];  // lg[0] < 0: 64 consecutive sites
        int dpw;    // directions per wavefront (4 / threads per site)
    };
    // tile order x tile shape at four threads per site, then two and one
    // thread per site on the best of those
    std::vector<cand> cands;
    const std::pair<int, unsigned> maps[] = {{0, 64}, {1, 64}, {2, 256}, {2, 1024}, {2, 4096}};
    const int shapes[][4] = {{-1, 0, 0, 0}, {2, 2, 1, 1}, {3, 2, 1, 0},
                             {3, 3, 0, 0},  {4, 2, 0, 0}, {2, 2, 2, 0}};
    for (const auto& sh : shapes) {
        qcb::lattice probe = h->g;
        if (sh[0] >= 0 && !qcb::set_tile_shape(probe, sh)) continue;
        if (sh[0] < 0) qcb::set_tile_shape(probe, nullptr);
        for (const auto& m : maps) cands.push_back({m.first, m.second, {sh[0], sh[1], sh[2], sh[3]}, 1});
        // slab sweeps: 8 to 128 slabs per t-slice (one to sixteen per XCD)
        const unsigned pt = qcb::stage_tiles_per_t(probe);
        for (unsigned ns : {8u, 16u, 32u, 64u, 128u})
            if (pt && pt % ns == 0) cands.push_back({3, pt / ns, {sh[0], sh[1], sh[2], sh[3]}, 1});
    }
    const size_t first_pass = cands.size();
    const qcb::lattice g0 = h->g;
    cand best = {g0.map_mode, g0.map_win, {g0.lg_tile[0], g0.lg_tile[1], g0.lg_tile[2], g0.lg_tile[3]},
                 g0.dpw};
    float best_ms = 1e30f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = QCBF_OK;
    if ((e = hipMemcpyAsync(saveU, h->U[cur0], ub, hipMemcpyDeviceToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(saveA, h->acc, ab, hipMemcpyDeviceToDevice, h->stream)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        rc = fail(h, QCBF_ERR_HIP, "autotune setup: %s", hipGetErrorString(e));
    for (size_t k = 0; rc == QCBF_OK && k < cands.size(); k++) {
        const cand c = cands[k];
        h->g.map_mode = c.mode;
        h->g.map_win = c.win;
        h->g.dpw = c.dpw;
        qcb::set_tile_shape(h->g, c.lg[0] < 0 ? nullptr : c.lg);
        rc = rk3_once(h, eps);  // warm
        if (rc) break;
        if ((e = hipEventRecord(e0, h->stream)) != hipSuccess) {
            rc = fail(h, QCBF_ERR_HIP, "event: %s", hipGetErrorString(e));
            break;
        }
        for (int s = 0; rc == QCBF_OK && s < trial_steps; s++) rc = rk3_once(h, eps);
        if (rc) break;
        float ms = 0.f;
        if ((e = hipEventRecord(e1, h->stream)) != hipSuccess ||
            (e = hipEventSynchronize(e1)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
            rc = fail(h, QCBF_ERR_HIP, "autotune timing: %s", hipGetErrorString(e));
            break;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
        if (k + 1 == first_pass)
            for (int d : {2, 4}) {
                cand t = best;
                t.dpw = d;
                cands.push_back(t);
            }
    }
    h->g.map_mode = best.mode;
    h->g.map_win = best.win;
    h->g.dpw = best.dpw;
    qcb::set_tile_shape(h->g, best.lg[0] < 0 ? nullptr : best.lg);
    h->cur = cur0;
    h->timing = timing0;
    h->rows_stale = stale0;
    if ((e = hipMemcpyAsync(h->U[cur0], saveU, ub, hipMemcpyDeviceToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->acc, saveA, ab, hipMemcpyDeviceToDevice, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->dwarn, &warn0, sizeof(int), hipMemcpyHostToDevice, h->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        if (rc == QCBF_OK) rc = fail(h, QCBF_ERR_HIP, "autotune restore: %s", hipGetErrorString(e));
    (void)hipFree(saveU);
    (void)hipFree(saveA);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (mode) *mode = best.mode;
    if (win) *win = best.win;
    return rc;
}
