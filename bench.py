"""Throughput of the Wilson-flow measurement path on one MI355X.

Workload (default): a 32^4 SU(3) lattice in FP64, Wilson flow at eps = 0.02
with the clover Q, E and the plaquette measured after every RK3 step.  One
bench "step" = one RK3 step (three fused stage kernels) + one measurement.
The timed region runs K steps through qcbf_flow_measure (K RK3 steps and
K+1 measurements, one synchronisation at the end).

Prints ONE JSON line:  value = site-flow-steps/s = V * K / seconds.
The roofline object reports the dominant kernel (the fused RK3 stage
kernel; its three instantiations are timed together with HIP events on the
handle's stream inside the timed region).  The CPU baseline is the FP64
oracle (a C restatement of the reference flow) on the host cores.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ecmc-qcb_amd"))

METRIC = "site·flow-steps/s, FP32-mixed vs FP64 GPU at |dE/E|<1e-6; HBM GB/s vs peak"
UNIT = "site·flow-steps/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic work per site per RK3 step (FP64 path):
#   bytes: each stage reads and writes the 4 links (144 B each) and the su(3)
#   accumulator (4 x 64 B) is written in stages 1-2 and read in stages 2-3:
#   4 * (3*2*144 + 4*64) = 4480 B.
#   flops: 4 links x 3 stages x 3371 = 40452.
FLOW_BYTES_PER_SITE_STEP = 4480.0
FLOW_FLOPS_PER_SITE_STEP = 40452.0
MEASURE_FLOPS_PER_SITE = 15150.0
# FP32-mixed form: U as 12 FP32 reals (48 B/link), accumulator 8 FP32 (32 B/link):
#   4 * (3*2*48 + 4*32) = 1664 B per site per RK3 step; +504 flops for the
#   third-row reconstruction.
MIXED_BYTES_PER_SITE_STEP = 1664.0
MIXED_FLOPS_PER_SITE_STEP = 40956.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(lattice):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary of
    this workload, or None."""
    tag = "x".join(str(v) for v in lattice)
    path = os.path.join(ROOT, "profiles", f"rk3_stage_f64_{tag}_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
    return float(m.group(1)) if m else None


def cpu_baseline(links, L, eps, steps, threads):
    """The FP64 oracle (C, OpenMP) on the same 32^4 field: `steps` RK3 steps,
    each followed by one measurement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.set_num_threads(threads)
    f = links.copy()
    secs = oracle.time_flow(f, L, eps, steps, measure=True)
    V = 1
    for v in L:
        V *= v
    return {
        "value": V * steps / secs,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} RK3 steps + {steps} measurements of the same "
                  f"{'x'.join(map(str, L))} FP64 field, {secs:.2f} s, "
                  f"OpenMP {threads} threads (C oracle, gcc -O2 -fopenmp)",
    }


def timed_flow(gf, K, warmup):
    """K bench steps (RK3 step + measurement) on an already filled handle.
    Returns (seconds, tQE, plaquettes, per-class kernel ms, per-class launches)."""
    if warmup > 0:
        gf.flow_measure(warmup + 1, 1)
    gf.synchronize()
    gf.timing(True)
    gf.synchronize()
    t0 = time.perf_counter()
    tqe, plaq = gf.flow_measure(K + 1, 1)
    gf.synchronize()
    secs = time.perf_counter() - t0
    ms, launches = gf.timing_read()
    gf.timing(False)
    return secs, tqe, plaq, ms, launches


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--L", type=int, nargs="+", default=[32])
    ap.add_argument("--eps", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-mixed leg")
    ap.add_argument("--map-mode", type=int, default=None,
                    help="workgroup->tile order (qcbf_tune); default: the library's")
    ap.add_argument("--map-win", type=int, default=None)
    ap.add_argument("--tile-shape", type=int, default=None, help="qcbf_tune tile-shape code")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip qcbf_autotune before the warmup")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run the plain copy kernel on this many MiB (traffic calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the whole lattice lives on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    L = tuple(args.L * 4) if len(args.L) == 1 else tuple(args.L)
    V = 1
    for v in L:
        V *= v
    import qcbflow

    name, cus = qcbflow.device_info()
    log(f"device {name}, {cus} CUs; lattice {L}, V = {V}")
    def setup(h):
        """launch shape: explicit knobs, then (unless disabled) the autotuner"""
        for key, val in ((qcbflow.TUNE_MAP_MODE, args.map_mode),
                         (qcbflow.TUNE_MAP_WIN, args.map_win),
                         (qcbflow.TUNE_TILE_SHAPE, args.tile_shape)):
            if val is not None:
                h.tune(key, val)
        if not args.no_autotune:
            h.autotune(trial_steps=2)
        return [h.get_tune(k) for k in (qcbflow.TUNE_MAP_MODE, qcbflow.TUNE_MAP_WIN,
                                        qcbflow.TUNE_TILE_SHAPE)]

    gf = qcbflow.GradientFlow(L, args.eps)
    gf.hot_start(args.seed)
    shape64 = setup(gf)
    log(f"fp64 launch shape (map mode, window, tile code): {shape64}")

    # timed region: K RK3 steps, K+1 measurements, stage kernels bracketed by events
    K = args.steps
    secs, tqe, plaq, ms, launches = timed_flow(gf, K, args.warmup)
    value = V * K / secs
    log(f"timed: {K} steps in {secs:.4f} s -> {value:.4e} site-steps/s; "
        f"E(t={tqe[-1, 0]:.2f}) = {tqe[-1, 2]:.6g}, Q = {tqe[-1, 1]:.6g}")

    stage_ms = ms[0] + ms[1] + ms[2]
    stage_n = launches[0] + launches[1] + launches[2]
    avg_stage_ms = stage_ms / stage_n
    bytes_per_launch = V * FLOW_BYTES_PER_SITE_STEP / 3.0
    achieved = bytes_per_launch / (avg_stage_ms * 1e-3) / 1e9
    per_stage = {f"stage{k + 1}_avg_ms": ms[k] / max(1, launches[k]) for k in range(3)}
    meas_avg_ms = ms[3] / max(1, launches[3])
    flops_rate = V * FLOW_FLOPS_PER_SITE_STEP / (3 * avg_stage_ms * 1e-3) / 1e12
    traffic = traffic_from_profiles(L)

    # flow-only throughput (no measurements)
    gf.synchronize()
    t2 = time.perf_counter()
    gf.rk3_step(K)
    gf.synchronize()
    flow_only = V * K / (time.perf_counter() - t2)

    copy_ms = None
    if args.calibrate_copy_mib > 0:
        copy_ms = gf.copy_bench(args.calibrate_copy_mib << 20, 10)
        log(f"copy kernel: {args.calibrate_copy_mib} MiB in {copy_ms:.4f} ms -> "
            f"{2 * (args.calibrate_copy_mib << 20) / (copy_ms * 1e-3) / 1e9:.1f} GB/s (read+write)")

    # the FP32-mixed form on the same lattice and cadence
    mixed = None
    if not args.no_mixed:
        links = gf.field() if not args.no_cpu else None
        gm = qcbflow.GradientFlow(L, args.eps, qcbflow.FP32_MIXED)
        gm.hot_start(args.seed)
        shape32 = setup(gm)
        msecs, mtqe, mplaq, mms, mn = timed_flow(gm, K, args.warmup)
        mstage = (mms[0] + mms[1] + mms[2]) / (mn[0] + mn[1] + mn[2])
        mach = V * MIXED_BYTES_PER_SITE_STEP / 3.0 / (mstage * 1e-3) / 1e9
        mixed = {
            "value": V * K / msecs,
            "vs_fp64": (V * K / msecs) / value,
            "ms_per_step": msecs / K * 1e3,
            "stage_avg_ms": mstage,
            "measure_avg_ms": mms[3] / max(1, mn[3]),
            "hbm_achieved_gbs": mach,
            "hbm_frac": mach / HBM_PEAK_GBS,
            "bytes_per_site_step": MIXED_BYTES_PER_SITE_STEP,
            "final_t": float(mtqe[-1, 0]),
            "dE_rel_vs_fp64_at_final_t": float(abs(mtqe[-1, 2] - tqe[-1, 2]) / tqe[-1, 2]),
            "dQ_vs_fp64_at_final_t": float(abs(mtqe[-1, 1] - tqe[-1, 1])),
            "launch_shape": shape32,
            "note": "same Haar hot start rounded to FP32; accuracy gates on thermalised "
                    "lattices are in tests/test_gpu_mixed.py",
        }
        log(f"fp32-mixed: {mixed['value']:.4e} site-steps/s ({mixed['vs_fp64']:.2f}x fp64), "
            f"stage {mstage:.4f} ms")
        gm.close()
    else:
        links = None

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        if links is None:
            links = gf.field()
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_steps} steps on {threads} threads ...")
        cpu = cpu_baseline(links, L, args.eps, args.cpu_steps, threads)
        log(f"cpu baseline: {cpu['value']:.4e} {UNIT}")
    gf.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": secs / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: Haar-random SU(3) hot start (seed {args.seed}) generated on the GPU; "
                "the flow's cost is data-independent (fixed 18-step Horner, no branches)",
        "config": {
            "workload": f"Wilson flow FP64, {'x'.join(map(str, L))} lattice, eps={args.eps}, "
                        "plaquette+Q+E measured every RK3 step",
            "lattice": list(L),
            "sites": V,
            "eps": args.eps,
            "measure_every_rk3_steps": 1,
            "precision": "fp64",
            "parallelism": "single GPU",
            "launch_shape": shape64,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_rk3_stage64<1,2,3> (fused RK3 stage)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_stage_ms,
            "fp64_tflops": flops_rate,
            "fp64_frac": flops_rate / FP64_PEAK_TFLOPS,
            **per_stage,
        },
        "measure_avg_ms": meas_avg_ms,
        "flow_only_value": flow_only,
        "fp32_mixed": mixed,
        "cpu_baseline": cpu,
    }
    if copy_ms is not None:
        out["copy_calibration"] = {"mib": args.calibrate_copy_mib, "avg_ms": copy_ms}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
