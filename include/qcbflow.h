/*
 * qcbflow.h -- C ABI of libqcbflow, the MI355X (gfx950) Wilson-flow
 * measurement library.
 *
 * Replaces the Wilson-flow measurement path of ECMC-QCB:
 *   class GradientFlow            src/flow/gradient_flow.h:11-27
 *   namespace mpi::observables    src/observables/observables_mpi.h:8-19
 *   read_ildg_clime / save_ildg_clime  src/io/ildg.h:17-20
 *   exp_su3_luscher               src/su3/utils.h:45
 * The C++ drop-in headers (ecmc-qcb_amd/compat/) keep those exact signatures
 * on top of this ABI.
 *
 * Conventions
 *  - Every function returns an int status (QCBF_OK = 0).  On failure
 *    qcbf_last_error(h) (or qcbf_last_error(NULL) for calls without a handle)
 *    returns a message.  Nothing throws across this boundary.
 *  - Host arrays of links use the reference's interior order, which is also
 *    the ILDG payload order: site s = ((t*Lz + z)*Ly + y)*Lx + x (x fastest),
 *    then mu = 0..3, then the 3x3 matrix row-major, each entry (re, im):
 *    72 doubles per site (src/io/ildg.cpp:233-246).
 *  - A handle owns all of its device memory and one HIP stream; calls on one
 *    handle are ordered on that stream.  A handle is not thread-safe: use one
 *    per host thread.  Host arrays are only read/written during the call.
 *  - Observables use the reference's conventions (observables_mpi.cpp):
 *      P = <Re tr U_p>/3 over the 6 planes and all sites,
 *      E = (1/V) sum_x 1/2 sum_{mu<nu} |F_munu(x)|_F^2,
 *      Q = (1/4pi^2) sum_x [Re tr F01F23 - Re tr F02F13 + Re tr F03F12],
 *    with F = (C - C^dag)/8 from the four-leaf clover C.
 */
#ifndef QCBFLOW_H
#define QCBFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QCBFLOW_ABI_VERSION 1

enum qcbf_status {
    QCBF_OK = 0,
    QCBF_ERR_ARG = 1,     /* invalid argument (null pointer, bad extent, ...) */
    QCBF_ERR_HIP = 2,     /* HIP runtime error */
    QCBF_ERR_IO = 3,      /* file cannot be opened / read / written */
    QCBF_ERR_FORMAT = 4,  /* malformed LIME/ILDG file or size mismatch */
    QCBF_ERR_OOM = 5,     /* device or pinned host allocation failed */
    QCBF_ERR_STATE = 6    /* call not valid in the handle's current state */
};

enum qcbf_precision {
    QCBF_FP64 = 0,       /* links and arithmetic in FP64 (plain path) */
    QCBF_FP32_MIXED = 1  /* two-row FP32 links, FP32 flow arithmetic, FP64
                            Cayley-Hamilton coefficients, projection and sums */
};

/* Kernel classes reported by qcbf_timing_read. */
enum qcbf_kernel_class {
    QCBF_K_STAGE1 = 0,  /* fused RK3 stage 1 */
    QCBF_K_STAGE2 = 1,  /* fused RK3 stage 2 */
    QCBF_K_STAGE3 = 2,  /* fused RK3 stage 3 */
    QCBF_K_MEASURE = 3, /* clover/plaquette/Q/E measurement + its reduction */
    QCBF_K_STAGE1_MEASURE = 4, /* stage 1 fused with the measurement of its input
                                  (qcbf_flow_measure), reduction included */
    QCBF_K_COUNT = 5
};

typedef struct qcbf_flow qcbf_flow;

/* Message of the last failed call on h (h == NULL: last failure on this thread
 * of a call made without a handle, or of qcbf_create). */
const char* qcbf_last_error(const qcbf_flow* h);

/* GradientFlow(double epsilon, const GaugeField&, const GeometryCB&)
 * (src/flow/gradient_flow.cpp:12-13) + GeometryCB(L) (src/geometry/GeometryCB.cpp:5-9).
 * dims = {Lx, Ly, Lz, Lt}, each >= 2, product < 2^28 sites.  precision is
 * QCBF_FP64 or QCBF_FP32_MIXED.  Allocates the link field (cold start), its RK3
 * companion buffers and the measurement buffers on the current HIP device.
 * Every entry point below works in both precisions; host arrays are always
 * FP64 (the FP32-mixed field is re-unitarised and rounded on upload and its
 * third row rebuilt in FP64 on download). */
int qcbf_create(const int dims[4], int precision, qcbf_flow** out);
int qcbf_destroy(qcbf_flow* h);
int qcbf_get_dims(const qcbf_flow* h, int dims[4], int* precision);

/* GradientFlow::copy(const GaugeField&) (src/flow/gradient_flow.cpp:138):
 * upload V*72 doubles in the reference order into the flowing field. */
int qcbf_set_links(qcbf_flow* h, const double* links);
/* Read the flowing field back (GradientFlow::field_c). */
int qcbf_get_links(qcbf_flow* h, double* links);

/* Copy the flowing field of src into dst on the device (same extents; the
 * precisions may differ: FP64 -> FP32-mixed re-unitarises and rounds, the
 * other way rebuilds the third row in FP64).  Both handles must be on the
 * same device. */
int qcbf_copy_field(qcbf_flow* dst, qcbf_flow* src);

/* GaugeField::hot_start / cold_start (src/gauge/GaugeField.cpp:10-33):
 * Haar-random SU(3) links from a counter-based generator keyed by seed, or the
 * identity.  Generated on the device. */
int qcbf_hot_start(qcbf_flow* h, uint64_t seed);
int qcbf_cold_start(qcbf_flow* h);

/* n_sweeps Cabibbo-Marinari heatbath sweeps at beta on the flowing field
 * (src/heatbath/heatbath.cpp:8-44, src/heatbath/heatbath_mpi.cpp:10-33): the
 * step that thermalises the flow's inputs, here on the device.  Checkerboard
 * by (mu, site parity), so the extents must be even; FP64 handles only
 * (QCBF_ERR_STATE otherwise).  Links are re-unitarised after every sweep.
 * Successive calls continue one random stream keyed by seed and the running
 * sweep count. */
int qcbf_heatbath(qcbf_flow* h, double beta, int n_sweeps, uint64_t seed);

/* mpi::shift::shift_field (src/mpi/Shift.cpp:109-160) on the whole periodic
 * lattice: U_mu(x) <- U_mu(x - shift) for every direction.  shift[d] is taken
 * modulo L[d] (negative values shift the other way).  Asynchronous, like
 * qcbf_rk3_step.  The reference's random_shift (Shift.cpp:163-186) draws the
 * shift on the host: the compat header mpi/Shift.h does that and calls this. */
int qcbf_shift(qcbf_flow* h, const int shift[4]);

/* read_ildg_clime (src/io/ildg.cpp:138-237): load an ILDG/LIME file straight
 * into the flowing field.  path is the file itself (the reference's
 * dirpath/filename/filename composition lives in the C++ drop-in).  The
 * ildg-format record, when present, must match the handle's extents;
 * precision 64 and 32 payloads are accepted; the ildg-binary-data size must
 * be V*72*precision/8 bytes.  Errors: QCBF_ERR_IO, QCBF_ERR_FORMAT. */
int qcbf_load_ildg(qcbf_flow* h, const char* path);
/* save_ildg_clime (src/io/ildg.cpp:34-136): write the flowing field as a
 * 64-bit big-endian ILDG file with the reference's record framing
 * (ildg-format MB=1 ME=1, then ildg-binary-data MB=0 ME=1). */
int qcbf_save_ildg(qcbf_flow* h, const char* path);

/* GradientFlow::rk3_step (src/flow/gradient_flow.cpp:124-135), n_steps times,
 * with flow step eps.  Three fused kernels per step; asynchronous on the
 * handle's stream. */
int qcbf_rk3_step(qcbf_flow* h, double eps, int n_steps);

/* The unfused GradientFlow member functions, for callers that drive the RK3
 * stages themselves (src/flow/gradient_flow.cpp:17-121).  Forces are kept as
 * su(3) fields of 8 reals per link. */
int qcbf_compute_force(qcbf_flow* h, int i);   /* i = 0 -> force_0, 1 -> force_1 */
int qcbf_compute_w1(qcbf_flow* h, double eps); /* U <- exp(eps/4 F0) U */
int qcbf_compute_w2(qcbf_flow* h, double eps); /* U <- exp(8/9 eps F1 - 17/36 eps F0) U */
int qcbf_replace_force_0(qcbf_flow* h);        /* F0 <- 8/9 F1 - 17/36 F0 */
int qcbf_compute_w3(qcbf_flow* h, double eps); /* U <- exp(3/4 eps F1 - eps F0) U */
/* Read force_i back as full 3x3 matrices in the link order (V*72 doubles). */
int qcbf_get_force(qcbf_flow* h, int i, double* out);

/* mean_plaquette_global (src/observables/observables_mpi.cpp:39-48). */
int qcbf_plaquette(qcbf_flow* h, double* plaq);
/* topo_q_e_clover_global (src/observables/observables_mpi.cpp:136-156):
 * Q and E of the flowing field. */
int qcbf_q_e_clover(qcbf_flow* h, double* Q, double* E);
/* Both in one fused pass. */
int qcbf_measure(qcbf_flow* h, double* plaq, double* E, double* Q);

/* clover_site (src/observables/observables_mpi.cpp:66-90) of the flowing field
 * at one site (site = ((t*Lz + z)*Ly + y)*Lx + x, 0 <= site < V):
 * out[18] = F_munu = (C - C^dag)/8 row-major (re, im), the trace kept; zero
 * for mu == nu (the reference prints a warning there). */
int qcbf_clover_site(qcbf_flow* h, long site, int mu, int nu, double* out);
/* local_q_e_clover (src/observables/observables_mpi.cpp:93-112) at one site:
 * q = (Re tr F01F23 - Re tr F02F13 + Re tr F03F12)/(4 pi^2),
 * e = 1/2 sum_{mu<nu} |F_munu|_F^2. */
int qcbf_local_q_e(qcbf_flow* h, long site, double* q, double* e);

/* topo_charge_flowed (src/observables/observables_mpi.cpp:159-186) on the
 * current flowing field: for n = 0..n_steps_gf-1 measure (t, Q, E) at t
 * (tQE[3n..3n+2]) and the plaquette (plaq[n], may be NULL); unless n is the
 * last, run n_rk_steps RK3 steps and t += eps after each.  Results stay on the
 * device until the end of the call (one synchronisation).  The flowing field
 * holds the flowed links afterwards. */
int qcbf_flow_measure(qcbf_flow* h, double eps, int n_steps_gf, int n_rk_steps, double* tQE,
                      double* plaq);

/* *flagged = 1 if any exponential since the previous call had an argument with
 * t = -1/2 tr X^2 > 1 (exp_su3_luscher warns on stderr in that case,
 * src/su3/utils.cpp:153-155), else 0.  Reading clears the flag. */
int qcbf_exp_warnings(qcbf_flow* h, int* flagged);

/* max_{links,ij} |(U U^dag - 1)_ij| and max |det U - 1| of the flowing field
 * (a NaN anywhere reports +inf). */
int qcbf_unitarity(qcbf_flow* h, double* max_dev, double* max_det_dev);

/* Wait for all work queued on the handle's stream. */
int qcbf_synchronize(qcbf_flow* h);

/* exp_su3_luscher(Z, coeff) (src/su3/utils.cpp:145-193) on the device for n
 * traceless anti-hermitian 3x3 matrices (18 doubles each, row-major re/im).
 * *flagged (may be NULL) is set to 1 if any had t > 1. */
int qcbf_su3_exp(const double* Z, int n, double coeff, double* out, int* flagged);

/* ---- allocation-free measurement (no counterpart in the reference) ------- */
/* The measurement of qcbf_measure (same kernels, same results bit for bit) on
 * a link field the caller already holds in device memory, for callers that
 * cannot allocate device memory: code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Needs no handle and allocates
 * nothing: its per-tile partial sums live in a fixed device array of the
 * library, 2^18 tiles (6 MiB per GPU, kept until the process ends), so the
 * volume is limited to 2^24 sites (64^4); larger volumes are rejected with
 * QCBF_ERR_ARG.  Calls that overlap on one GPU share that array and race.
 *   links: the field in the library's device layout for `precision`, with
 *          V = Lx*Ly*Lz*Lt and site s = ((t*Lz + z)*Ly + y)*Lx + x:
 *          QCBF_FP64: entry k = 3i+j of U_mu(s) as (re, im) doubles at
 *            double2 index (mu*9 + k)*V + s (rows 0 and 1 are read; the
 *            third row is rebuilt as conj(row0 x row1));
 *          QCBF_FP32_MIXED: rows 0-1 as three float4 per link at float4
 *            index (mu*3 + w)*V + s, w = 0: (U00, U01), 1: (U02, U10),
 *            2: (U11, U12), each complex as (re, im).
 *   out:   device memory for 3 doubles, {P, E, Q} in the conventions above.
 *   stream: a hipStream_t (NULL: the default stream); the call is
 *          asynchronous on it.
 * Errors are read with qcbf_last_error(NULL). */
int qcbf_measure_device(const void* links, const int dims[4], int precision, double* out,
                        void* stream);

/* ---- launch tuning (no counterpart in the reference) --------------------- */
enum qcbf_tune_key {
    QCBF_TUNE_MAP_MODE = 0, /* workgroup -> site-tile order of the flow kernels:
                               0 linear, 1 one contiguous chunk per XCD,
                               2 XCD chunks inside windows of MAP_WIN workgroups,
                               3 slab sweep: each XCD walks slabs of MAP_WIN tiles
                                 of a t-slice through t (single lattices only) */
    QCBF_TUNE_MAP_WIN = 1,  /* mode 2: window (rounded down to a multiple of 8, at
                               least 8); mode 3: slab width in tiles; >= 1.  With
                               fewer than 8 slabs per t-slice some XCDs get no
                               slab (32^4: MAP_WIN 128 leaves half of them idle).
                               The default takes 16 slabs per t-slice, more when
                               that keeps slabs at most 256 tiles (96^4: 64);
                               qcbf_autotune tries 8 to 128 */
    QCBF_TUNE_BLOCKS_PER_CU = 2, /* ensembles only: persistent-grid workgroups per CU
                                    (1..16), or 0 (default) for one workgroup per tile */
    QCBF_TUNE_FUSED_MEASURE = 4, /* qcbf_flow_measure fuses each measurement into the
                                    next RK3 stage 1: 1 (default) or 0 (separate
                                    kernels); results are bit-identical either way.
                                    Needs the default tile shape. */
    QCBF_TUNE_TILE_SHAPE = 3,    /* sites of one flow workgroup: -1 = 64 consecutive
                                    sites, else a 4-D box of extents 2^a x 2^b x 2^c x 2^d
                                    (a+b+c+d = 6) encoded a + 8b + 64c + 512d; the box
                                    must divide the lattice (ensembles: applied to each
                                    member it divides) */
    QCBF_TUNE_THREADS_PER_SITE = 5, /* threads per site of the RK3 stage kernels: 4
                                      (default; one per direction), 2 or 1 (each thread
                                      updates 2 or 4 directions in turn, and a workgroup
                                      takes 2 or 4 tiles).  The fused measurement stage
                                      keeps 4. */
    QCBF_TUNE_LDS_STAGE = 6      /* 1 (default) runs the RK3 stages with the links the
                                    four direction wavefronts of a tile share staged in
                                    LDS, 0 the plain stage kernel; same results bit for
                                    bit.  Applies with 64-consecutive-site tiles, 4
                                    threads per site and a volume divisible by 64, for
                                    ensembles every member's, on the one-workgroup-per-
                                    tile grid (the plain kernel runs otherwise). */
};
/* Set one launch parameter of h; results do not depend on it (only the order in
 * which workgroups visit sites changes).  QCBF_ERR_ARG for an unknown key or
 * value. */
int qcbf_tune(qcbf_flow* h, int key, long value);
int qcbf_get_tune(const qcbf_flow* h, int key, long* value);

/* Time each candidate launch shape of the flow kernels (tile order x 4-D tile
 * shape at 4 threads per site, then 2 and 1 thread per site on the fastest of
 * those; see qcbf_tune_key) for trial_steps RK3 steps on the real field, which
 * is saved first and restored afterwards, and keep the fastest; *mode and *win
 * (may be NULL) receive the chosen order, qcbf_get_tune the tile shape and the
 * threads per site. */
int qcbf_autotune(qcbf_flow* h, double eps, int trial_steps, long* mode, long* win);

/* ---- ensembles (no counterpart in the reference: its callers flow one
 * configuration at a time through topo_charge_flowed) ---------------------- */
/* n configurations of extents dims[4c..4c+3] (each >= 2, volume < 2^28), one
 * precision.  All fields live in one allocation; one persistent kernel per RK3
 * stage walks the (configuration, 64-site tile) list of all of them. */
typedef struct qcbf_batch qcbf_batch;
const char* qcbf_batch_last_error(const qcbf_batch* b);
int qcbf_batch_create(int n, const int* dims, int precision, qcbf_batch** out);
int qcbf_batch_destroy(qcbf_batch* b);
int qcbf_batch_size(const qcbf_batch* b, int* n, int* dims);
int qcbf_batch_set_links(qcbf_batch* b, int c, const double* links);
int qcbf_batch_get_links(qcbf_batch* b, int c, double* links);
int qcbf_batch_hot_start(qcbf_batch* b, int c, uint64_t seed);
/* Configuration c <- the flowing field of src (same extents), on the device. */
int qcbf_batch_copy_from(qcbf_batch* b, int c, qcbf_flow* src);
int qcbf_batch_rk3_step(qcbf_batch* b, double eps, int n_steps);
/* {P, E, Q} of every configuration: out[3c..3c+2] */
int qcbf_batch_measure(qcbf_batch* b, double* out);
/* topo_charge_flowed cadence for every configuration at once:
 * tQE[(c*n_steps_gf + n)*3 + {0,1,2}] = (t, Q, E), plaq[c*n_steps_gf + n]
 * (plaq may be NULL).  With graphs on and n_rk_steps = 1, each (measure, RK3
 * step) pair is one replay of a captured hipGraph. */
int qcbf_batch_flow_measure(qcbf_batch* b, double eps, int n_steps_gf, int n_rk_steps,
                            double* tQE, double* plaq);
int qcbf_batch_exp_warnings(qcbf_batch* b, int* flagged);
int qcbf_batch_synchronize(qcbf_batch* b);
int qcbf_batch_tune(qcbf_batch* b, int key, long value);
int qcbf_batch_get_tune(const qcbf_batch* b, int key, long* value);
int qcbf_batch_use_graph(qcbf_batch* b, int on);
/* Autotuner: tile order x persistent-grid size x tile shape at 4 threads per
 * site, then 2 and 1 thread per site on the fastest of those, timed on the
 * real fields (saved first, restored afterwards); the fastest is kept. */
int qcbf_batch_autotune(qcbf_batch* b, double eps, int trial_steps, long* mode, long* win,
                        long* blocks_per_cu);
/* Returns the device time of the flow_measure calls since the previous call
 * (total ms and number of RK3 steps), then switches the timer on or off. */
int qcbf_batch_timing(qcbf_batch* b, int on, double* total_ms, long* rk3_steps);

/* ---- measurement support (no counterpart in the reference) -------------- */
/* With on = n >= 1, every n-th launch of each kernel class on h (the first,
 * the (n+1)-th, ...) is bracketed by HIP events on h's stream; on = 0 stops.
 * n = 1 times every launch.  A larger n keeps most of the event markers, which
 * add a few microseconds between kernels, out of a timed loop.
 * qcbf_timing_read synchronises and returns, per kernel class, the summed
 * milliseconds and the number of timed launches since timing was
 * (re)enabled. */
int qcbf_timing_enable(qcbf_flow* h, int on);
int qcbf_timing_read(qcbf_flow* h, double total_ms[QCBF_K_COUNT], long launches[QCBF_K_COUNT]);
/* Device-to-device copy of `bytes` through a plain copy kernel, repeated reps
 * times; *ms = average kernel milliseconds (HIP events). */
int qcbf_copy_bench(qcbf_flow* h, size_t bytes, int reps, double* ms);
/* Name and compute-unit count of the current device. */
int qcbf_device_info(char* name, int name_len, int* compute_units);

#ifdef __cplusplus
}
#endif
#endif /* QCBFLOW_H */
