// noalloc.hip -- qcbf_measure_device: the clover / plaquette / Q / E
// measurement of a link field the caller already holds on the device, for
// callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// It runs the kernels of qcbf_measure (measure.hip: the per-tile partial
// triples, then their ordered, compensated reduction) with the partials in
// `scratch`, a file-scope device array of kNoAllocMaxTiles triples, instead
// of a handle's buffer; nothing is allocated.  Every call on one GPU uses
// that GPU's copy of the array, so calls that overlap on one GPU race; each
// call writes every partial it reads, so nothing needs clearing between
// calls.
#include <mutex>

#include "api_internal.hpp"
#include "kernels.hpp"
#include "qcbflow.h"

// Up to 2^18 tiles of 64 sites: volumes up to 2^24 sites (64^4), 6 MiB.
constexpr unsigned kNoAllocMaxTiles = 1u << 18;
static __device__ __align__(256) double scratch[3 * kNoAllocMaxTiles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// the message is read with qcbf_last_error(NULL)
int fail(int code, const char* msg) { return qcb::fail_without_handle(code, msg); }

}  // namespace

extern "C" int qcbf_measure_device(const void* links, const int dims[4], int precision,
                                   double* out, void* stream) {
    if (!links || !dims || !out) return fail(QCBF_ERR_ARG, "qcbf_measure_device: null argument");
    if (precision != QCBF_FP64 && precision != QCBF_FP32_MIXED)
        return fail(QCBF_ERR_ARG, "qcbf_measure_device: unknown precision");
    unsigned long long V = 1;
    for (int d = 0; d < 4; d++) {
        if (dims[d] < 2 || dims[d] > 4096)
            return fail(QCBF_ERR_ARG, "qcbf_measure_device: extent out of range [2, 4096]");
        V *= (unsigned long long)dims[d];
    }
    if (V > (1ull << 24))
        return fail(QCBF_ERR_ARG,
                    "qcbf_measure_device: more than 2^24 sites (its scratch holds 2^18 tiles)");
    const qcb::lattice g = qcb::make_lattice(dims);
    if (qcb::measure_blocks(g) > kNoAllocMaxTiles)
        return fail(QCBF_ERR_ARG,
                    "qcbf_measure_device: more than 2^18 tiles (its scratch holds 2^18 tiles)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(QCBF_ERR_HIP, "qcbf_measure_device: no usable current device");
    double* partials = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_lock);
        if (!g_scratch[dev]) {
            // the device address of this device's copy of `scratch`
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                (void)hipGetLastError();
                return fail(QCBF_ERR_HIP, "qcbf_measure_device: hipGetSymbolAddress(scratch) failed");
            }
            g_scratch[dev] = static_cast<double*>(p);
        }
        partials = g_scratch[dev];
    }
    const qcb::prec p = precision == QCBF_FP64 ? qcb::prec::f64 : qcb::prec::f32m;
    const hipError_t e = qcb::launch_measure(p, links, g, partials, out, true,
                                             static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(QCBF_ERR_HIP, "qcbf_measure_device: kernel launch failed");
    }
    return QCBF_OK;
}
