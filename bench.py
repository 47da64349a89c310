"""Throughput of the Wilson-flow measurement path on one MI355X.

Workload (default): the large single lattice, 64^4 SU(3) (16.8M sites), Haar
hot start from seed 2026 thermalised by 50 checkerboard heatbath sweeps at
beta = 6.4 on the GPU, then Wilson flow at eps = 0.02 with the clover Q, E
and the plaquette measured every 10 RK3 steps.  One bench "step" is one RK3
step (three stage kernels); the timed region runs K steps (default 400)
through qcbf_flow_measure (K/10 + 1 measurements, one synchronisation at the
end).  `--preset c2` runs the 32^4 lattice (beta 6.0, seed 7, 50 sweeps)
measured after every step instead.

Prints ONE JSON line:  value = FP64 site-flow-steps/s = V * K / seconds.
Beside it: the FP32-mixed form on the same input and cadence (`fp32_mixed`),
its accuracy against FP64 at flow time t = 2 from a separate untimed leg of
2/eps steps in both precisions (whatever K is), the roofline object of the
dominant kernel (the plain RK3 stage kernel, every launch of the timed
region bracketed by HIP events on the handle's stream), the batched form on
its own workload (`batched_ensemble`: the 64 lattices of uneven volumes of
`--report ensemble`, about 10 s, with the FP32-mixed batched flow against the
plain FP64 path flowing the same lattices one at a time) and the CPU baseline
(the FP64 oracle, a C restatement of the reference flow, on the host cores:
one RK3 step and one measurement of the same field, extrapolated).

--report precision | ensemble  runs one of the other workloads instead
(48^4 FP32-mixed vs FP64 accuracy and t0; the 64-lattice ensemble of uneven
volumes) and prints one JSON line for it.
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ecmc-qcb_amd"))

METRIC = "site·flow-steps/s, FP32-mixed vs FP64 GPU at |dE/E|<1e-6; HBM GB/s vs peak"
UNIT = "site·flow-steps/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# MI355X's published FP32 vector peak counts packed (two-wide) FP32 issue; this
# build compiles with packed FP32 switched off, which halves the issue rate.
FP32_PEAK_TFLOPS = 157.3
FP32_UNPACKED_TFLOPS = 78.6

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
# per stage (1, 2, 3): links read + written, accumulator written / read+written / read
STAGE_BYTES = {"fp64": (1408.0, 1664.0, 1408.0), "fp32_mixed": (512.0, 640.0, 512.0)}
# one read of the links by the measurement
MEASURE_BYTES = {"fp64": 576.0, "fp32_mixed": 192.0}

# Single-lattice workloads: (lattice extent, beta, seed, heatbath sweeps,
# RK3 steps between measurements, default timed steps).
PRESETS = {
    "large": (64, 6.4, 2026, 50, 10, 400),  # 64^4, measured every 10 steps
    "c2": (32, 6.0, 7, 50, 1, 200),         # 32^4, measured every step
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(lattice):
    """HBM bytes per FP64 stage launch from the committed rocprofv3 PMC summary
    of this workload (profiles/rk3_stage_f64_<lattice>_traffic.md or
    profiles/rk3_stage_<lattice>_kernel_trace_traffic.md), or None."""
    tag = "x".join(str(v) for v in lattice)
    for name in (f"rk3_stage_f64_{tag}_traffic.md", f"rk3_stage_{tag}_kernel_trace_traffic.md"):
        path = os.path.join(ROOT, "profiles", name)
        if os.path.exists(path):
            with open(path) as f:
                m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
            if m:
                return float(m.group(1))
    return None


def host_cpus():
    """CPUs this process may run on (nproc) and the host CPU model (lscpu's
    'Model name', from /proc/cpuinfo)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(links, L, eps, every, threads, steps=1):
    """The FP64 oracle (C, OpenMP) on the same field: `steps` RK3 steps and one
    measurement, each timed, combined at the workload's cadence (one
    measurement every `every` steps) into site-flow-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.set_num_threads(threads)
    f = links.copy()
    secs_flow = oracle.time_flow(f, L, eps, steps, measure=False)
    t0 = time.perf_counter()
    oracle.q_e_clover(f, L)
    oracle.plaquette(f, L)
    secs_meas = time.perf_counter() - t0
    del f
    V = volume(L)
    per_step = secs_flow / steps + secs_meas / every
    nproc, model = host_cpus()
    return {
        "value": V / per_step,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} RK3 step(s) ({secs_flow:.2f} s) and 1 measurement ({secs_meas:.2f} s) "
                  f"of the same {'x'.join(map(str, L))} FP64 field, extrapolated to the "
                  f"workload's cadence (1 measurement per {every} steps); OpenMP {threads} "
                  f"threads, nproc {nproc} ({model}); C oracle, gcc -O2 -fopenmp",
        "rk3_step_s": secs_flow / steps,
        "measure_s": secs_meas,
        "nproc": nproc,
        "cpu_model": model,
    }


def timed_flow(h, src, K, warmup, every=1, time_every=1):
    """Warmup on the handle, then refill it from `src` (a handle holding the
    input field) so the timed flow starts at t = 0 on the input, and run
    K RK3 steps with a measurement every `every` steps (K divisible by it).
    Returns (seconds, tQE, plaquettes, per-class kernel ms, per-class launches);
    the kernel times cover every `time_every`-th launch of each class."""
    if warmup > 0:
        h.flow_measure(warmup + 1, 1)
    h.copy_from(src)
    h.synchronize()
    h.timing(time_every)
    h.synchronize()
    t0 = time.perf_counter()
    tqe, plaq = h.flow_measure(K // every + 1, every)
    h.synchronize()
    secs = time.perf_counter() - t0
    ms, launches = h.timing_read()
    h.timing(False)
    return secs, tqe, plaq, ms, launches


def thermalised(qcbflow, L, beta, seed, sweeps):
    """FP64 handle: Haar hot start from `seed`, then `sweeps` heatbath sweeps."""
    g = qcbflow.GradientFlow(L)
    g.hot_start(seed)
    if sweeps > 0:
        g.heatbath(beta, sweeps, seed)
    g.synchronize()
    return g


def t_at(tqe, target):
    """Flow time where t^2 E(t) first reaches `target` (linear interpolation
    between measurements), or None when the flow stops short of it."""
    t, E = tqe[:, 0], tqe[:, 2]
    y = t * t * E
    k = np.nonzero(y >= target)[0]
    if len(k) == 0 or k[0] == 0:
        return None
    k = k[0]
    return float(t[k - 1] + (target - y[k - 1]) * (t[k] - t[k - 1]) / (y[k] - y[k - 1]))


def compare(t64, p64, t32, p32, t_gate=2.0):
    """Errors of the FP32-mixed series against the FP64 series (same input,
    same cadence): at flow time t_gate and maxima over the whole flow.  The
    gates are None unless the series holds a measurement at t_gate."""
    k = int(np.argmin(np.abs(t64[:, 0] - t_gate)))
    at_gate = abs(t64[k, 0] - t_gate) < 1e-9
    dE = np.abs(t32[:, 2] - t64[:, 2]) / np.abs(t64[:, 2])
    dQ = np.abs(t32[:, 1] - t64[:, 1])
    dP = np.abs(p32 - p64) / np.abs(p64)
    return {
        "t_gate": float(t64[k, 0]) if at_gate else None,
        "dE_rel_at_t_gate": float(dE[k]) if at_gate else None,
        "dQ_abs_at_t_gate": float(dQ[k]) if at_gate else None,
        "dP_rel_at_t_gate": float(dP[k]) if at_gate else None,
        "E_fp64_at_t_gate": float(t64[k, 2]) if at_gate else None,
        "Q_fp64_at_t_gate": float(t64[k, 1]) if at_gate else None,
        "max_dE_rel": float(dE.max()),
        "max_dQ_abs": float(dQ.max()),
        "max_dP_rel": float(dP.max()),
        "gate_dE_lt_1e-6": bool(dE[k] < 1e-6) if at_gate else None,
        "gate_dQ_lt_1e-4": bool(dQ[k] < 1e-4) if at_gate else None,
    }


def accuracy_leg(qcbflow, src, L, eps, every, t_end=2.0):
    """FP64 and FP32-mixed flows of the same input to t = t_end (t_end/eps RK3
    steps, measured every `every` steps, untimed); the FP32-mixed errors
    against FP64 at t_end, and where t^2 E first reaches 0.15 and 0.3."""
    n = int(round(t_end / eps))
    if n % every:
        every = math.gcd(n, every)
    series = {}
    for prec in (qcbflow.FP64, qcbflow.FP32_MIXED):
        with qcbflow.GradientFlow(L, eps, prec) as h:
            h.copy_from(src)
            series[prec] = h.flow_measure(n // every + 1, every)
    (t64, p64), (t32, p32) = series[qcbflow.FP64], series[qcbflow.FP32_MIXED]
    acc = compare(t64, p64, t32, p32, t_gate=n * eps)
    acc["rk3_steps"] = n
    acc["measured_every"] = every
    # E_ref is the reference's (1/2) sum |F|^2 / V, half the usual
    # normalisation: t^2 E_ref = 0.15 is the customary t0, 0.3 the literal one
    t0s = {}
    for thr in (0.15, 0.3):
        a, b = t_at(t64, thr), t_at(t32, thr)
        t0s[f"{thr}"] = {"fp64": a, "fp32_mixed": b,
                         "diff": None if a is None or b is None else abs(a - b)}
    acc["t_at_t2E"] = t0s
    return acc


def stage_roofline(ms, launches, V, tag):
    """Average duration (ms), algorithmic GB/s and average algorithmic bytes of
    the plain stage kernels (timing classes 0-2; a stage 1 fused with the
    measurement is class 4), bytes weighted by stage."""
    sb = STAGE_BYTES[tag]
    n = launches[0] + launches[1] + launches[2]
    t = ms[0] + ms[1] + ms[2]
    byt = V * sum(sb[k] * launches[k] for k in range(3))
    avg = t / max(1, n)
    return avg, (byt / (t * 1e-3) / 1e9 if t > 0 else 0.0), byt / max(1, n)


def fused_stats(ms, launches, V, tag):
    """The fused stage 1 + measurement kernel (timing class 4), or None."""
    if launches[4] == 0:
        return None
    avg = ms[4] / launches[4]
    byt = V * (STAGE_BYTES[tag][0] + MEASURE_BYTES[tag])
    return {"avg_ms": avg, "launches": launches[4], "bytes_per_launch": byt,
            "achieved_gbs": byt / (avg * 1e-3) / 1e9}


def setup_shape(qcbflow, h, args):
    """launch shape: the library's default (slab sweep over 64-site tiles),
    explicit knobs, then the autotuner if asked for"""
    for key, val in ((qcbflow.TUNE_MAP_MODE, args.map_mode),
                     (qcbflow.TUNE_MAP_WIN, args.map_win),
                     (qcbflow.TUNE_TILE_SHAPE, args.tile_shape),
                     (qcbflow.TUNE_THREADS_PER_SITE, args.threads_per_site)):
        if val is not None:
            h.tune(key, val)
    if args.lds_stage is not None:
        h.tune(qcbflow.TUNE_LDS_STAGE, args.lds_stage)
    if args.autotune:
        h.autotune(trial_steps=2)
    if args.no_fuse:
        h.tune(qcbflow.TUNE_FUSED_MEASURE, 0)
    return [h.get_tune(k) for k in (qcbflow.TUNE_MAP_MODE, qcbflow.TUNE_MAP_WIN,
                                    qcbflow.TUNE_TILE_SHAPE, qcbflow.TUNE_FUSED_MEASURE,
                                    qcbflow.TUNE_THREADS_PER_SITE, qcbflow.TUNE_LDS_STAGE)]


def lattice_of(args, default):
    if args.L is None:
        return tuple(default) if not np.isscalar(default) else (default,) * 4
    return tuple(args.L * 4) if len(args.L) == 1 else tuple(args.L)


def volume(L):
    V = 1
    for v in L:
        V *= v
    return V


# Device bytes per site a single-lattice handle keeps resident: two link
# fields (ping-pong) and the 8-real su(3) accumulator.  An FP32-mixed link is
# two rows (12 FP32 reals, 48 B); the 18-real FP32 form would take 72 B.
LINK_FIELD_BYTES_PER_SITE = {"fp64": 4 * 144, "fp32_mixed": 4 * 48}
ACC_BYTES_PER_SITE = {"fp64": 4 * 64, "fp32_mixed": 4 * 32}


def device_memory(V, tag):
    link = LINK_FIELD_BYTES_PER_SITE[tag] * V
    acc = ACC_BYTES_PER_SITE[tag] * V
    return {"link_field_gb": link / 1e9, "resident_gb": (2 * link + acc) / 1e9,
            "fp32_18_real_field_gb": 4 * 72 * V / 1e9 if tag == "fp32_mixed" else None}


def report_precision(qcbflow, args):
    """48^4 at beta = 6.2 from seed 11: FP64 and FP32-mixed flows of 250 RK3
    steps measured after every step; throughput, stage roofline, and the
    FP32-mixed errors against FP64 (P, E, Q at t = 2, t0)."""
    L = lattice_of(args, 48)
    K = args.steps if args.steps_given else 250
    beta, seed = 6.2, 11
    sweeps = args.sweeps if args.sweeps is not None else 100
    V = volume(L)
    log(f"precision: lattice {L}, V = {V}, beta {beta}, seed {seed}, {sweeps} heatbath sweeps")
    t0 = time.perf_counter()
    src = thermalised(qcbflow, L, beta, seed, sweeps)
    log(f"  thermalised in {time.perf_counter() - t0:.2f} s, plaquette {src.plaquette():.6f}")
    legs = {}
    series = {}
    for prec, tag, bps in ((qcbflow.FP64, "fp64", FLOW_BYTES_PER_SITE_STEP),
                           (qcbflow.FP32_MIXED, "fp32_mixed", MIXED_BYTES_PER_SITE_STEP)):
        h = qcbflow.GradientFlow(L, args.eps, prec)
        h.copy_from(src)
        shape = setup_shape(qcbflow, h, args)
        secs, tqe, plaq, ms, n = timed_flow(h, src, K, args.warmup, 1)
        avg, gbs, _ = stage_roofline(ms, n, V, tag)
        legs[tag] = {
            "value": V * K / secs,
            "ms_per_rk3_step": secs / K * 1e3,
            "stage_avg_ms": avg,
            "hbm_achieved_gbs": gbs,
            "hbm_frac": gbs / HBM_PEAK_GBS,
            "bytes_per_site_step": bps,
            "measure_avg_ms": ms[3] / max(1, n[3]),
            "stage1_measure": fused_stats(ms, n, V, tag),
            "launch_shape": shape,
            "exp_warnings": h.exp_warnings(),
            "device_memory": device_memory(V, tag),
        }
        series[tag] = (tqe, plaq)
        log(f"  {tag}: {legs[tag]['value']:.4e} site-steps/s, stage {avg:.4f} ms "
            f"({gbs:.0f} GB/s), E(t={tqe[-1, 0]:.2f}) = {tqe[-1, 2]:.6g}, Q = {tqe[-1, 1]:.6g}")
        h.close()
    src.close()
    t64, p64 = series["fp64"]
    t32, p32 = series["fp32_mixed"]
    acc = compare(t64, p64, t32, p32)
    t0s = {}
    for thr in (0.05, 0.1, 0.15, 0.3):
        a, b = t_at(t64, thr), t_at(t32, thr)
        t0s[f"{thr}"] = {"fp64": a, "fp32_mixed": b,
                         "diff": None if a is None or b is None else abs(a - b)}
    acc["t_at_t2E"] = t0s
    if acc["t_gate"] is not None:
        log(f"  fp32-mixed vs fp64 at t={acc['t_gate']:.2f}: dE/E {acc['dE_rel_at_t_gate']:.2e}, "
            f"dQ {acc['dQ_abs_at_t_gate']:.2e}, dP/P {acc['dP_rel_at_t_gate']:.2e}")
    return {
        "report": "precision",
        "metric": METRIC,
        "unit": UNIT,
        "config": {"workload": f"Wilson flow, {'x'.join(map(str, L))}, eps={args.eps}, "
                               f"{K} RK3 steps, measured every step",
                   "lattice": list(L), "sites": V, "beta": beta, "seed": seed,
                   "heatbath_sweeps": sweeps},
        "fp64": legs["fp64"],
        "fp32_mixed": legs["fp32_mixed"],
        "fp32_mixed_vs_fp64": legs["fp32_mixed"]["value"] / legs["fp64"]["value"],
        "accuracy_fp32_mixed_vs_fp64": acc,
        "data": "synthetic: GPU Haar hot start + GPU heatbath",
    }


ENSEMBLE_SIZES = (8, 12, 16, 24)
ENSEMBLE_BETAS = (5.8, 6.0, 6.2)


def report_ensemble(qcbflow, args):
    """64 configurations, 16 each of 8^4, 12^4, 16^4, 24^4, beta cycling over
    {5.8, 6.0, 6.2}, Haar hot starts from seeds 100..163 thermalised by 30
    heatbath sweeps, flowed 150 RK3 steps together (persistent stage kernels,
    graph-replayed steps, autotuned launch shape) in FP32-mixed and FP64,
    against the same configurations flowed one at a time."""
    K = args.steps if args.steps_given else 150
    sweeps = args.sweeps if args.sweeps is not None else 30
    dims = [(ENSEMBLE_SIZES[c // 16],) * 4 for c in range(64)]
    betas = [ENSEMBLE_BETAS[c % 3] for c in range(64)]
    seeds = [100 + c for c in range(64)]
    total_v = sum(volume(L) for L in dims)
    log(f"ensemble: 64 configurations, sum of volumes {total_v}, {K} steps, {sweeps} sweeps")
    ens = {p: qcbflow.Ensemble(dims, args.eps, p) for p in (qcbflow.FP64, qcbflow.FP32_MIXED)}
    seq = {qcbflow.FP64: [], qcbflow.FP32_MIXED: []}
    seq_secs = {qcbflow.FP64: 0.0, qcbflow.FP32_MIXED: 0.0}
    singles = {}
    t_start = time.perf_counter()
    for c in range(64):
        L = dims[c]
        src = thermalised(qcbflow, L, betas[c], seeds[c], sweeps)
        for p in ens:
            ens[p].copy_from(c, src)
            if (L, p) not in singles:
                singles[(L, p)] = qcbflow.GradientFlow(L, args.eps, p)
                singles[(L, p)].copy_from(src)
                singles[(L, p)].flow_measure(3, 1)  # code objects loaded before timing
            h = singles[(L, p)]
            h.copy_from(src)
            h.synchronize()
            t0 = time.perf_counter()
            tqe, plaq = h.flow_measure(K + 1, 1)
            h.synchronize()
            seq_secs[p] += time.perf_counter() - t0
            seq[p].append((tqe, plaq))
        src.close()
    log(f"  thermalised + sequential flows in {time.perf_counter() - t_start:.1f} s")
    for h in singles.values():
        h.close()
    out = {"report": "ensemble", "metric": METRIC, "unit": UNIT,
           "config": {"workload": "ensemble of 64 lattices (16 each of 8^4, 12^4, 16^4, 24^4), "
                                  f"eps={args.eps}, {K} RK3 steps, measured every step",
                      "sizes": list(ENSEMBLE_SIZES), "betas": list(ENSEMBLE_BETAS),
                      "seeds": [seeds[0], seeds[-1]], "heatbath_sweeps": sweeps,
                      "sum_sites": total_v},
           "data": "synthetic: GPU Haar hot starts + GPU heatbath"}
    res = {}
    for p, tag in ((qcbflow.FP64, "fp64"), (qcbflow.FP32_MIXED, "fp32_mixed")):
        e = ens[p]
        e.use_graph(True)
        if args.no_fuse:
            e.tune(qcbflow.TUNE_FUSED_MEASURE, 0)
        m, w, b = (None, None, None)
        if args.threads_per_site is not None:
            e.tune(qcbflow.TUNE_THREADS_PER_SITE, args.threads_per_site)
        if args.lds_stage is not None:
            e.tune(qcbflow.TUNE_LDS_STAGE, args.lds_stage)
        if not args.no_autotune:  # the ensemble configuration is autotuned
            m, w, b = e.autotune(trial_steps=2)
        tps = e.get_tune(qcbflow.TUNE_THREADS_PER_SITE)
        shape_code = e.get_tune(qcbflow.TUNE_TILE_SHAPE)
        e.synchronize()
        t0 = time.perf_counter()
        tqe, plaq = e.flow_measure(K + 1, 1)
        e.synchronize()
        secs = time.perf_counter() - t0
        same = all(np.array_equal(tqe[c], seq[p][c][0]) and np.array_equal(plaq[c], seq[p][c][1])
                   for c in range(64))
        res[tag] = (tqe, plaq)
        out[tag] = {
            "value": total_v * K / secs,
            "seconds": secs,
            "sequential_value": total_v * K / seq_secs[p],
            "speedup_vs_sequential": seq_secs[p] / secs,
            "launch_shape": {"map_mode": m, "map_win": w, "blocks_per_cu": b,
                             "tile_shape": shape_code, "threads_per_site": tps},
            "bit_identical_to_single_lattice_path": bool(same),
            "exp_warnings": e.exp_warnings(),
        }
        log(f"  {tag}: ensemble {out[tag]['value']:.4e} site-steps/s, sequential "
            f"{out[tag]['sequential_value']:.4e} ({out[tag]['speedup_vs_sequential']:.2f}x), "
            f"identical to single-lattice path: {same}")
        e.close()
    out["fp32_mixed_vs_fp64"] = out["fp32_mixed"]["value"] / out["fp64"]["value"]
    # the FP32-mixed batched flow against the plain FP64 path (the same 64
    # configurations flowed one at a time by the single-lattice FP64 kernels)
    out["fp32_mixed_batched_vs_fp64_plain"] = (out["fp32_mixed"]["value"] /
                                               out["fp64"]["sequential_value"])
    log(f"  fp32-mixed batched vs fp64 plain (one lattice at a time): "
        f"{out['fp32_mixed_batched_vs_fp64_plain']:.2f}x")
    per = [compare(res["fp64"][0][c], res["fp64"][1][c], res["fp32_mixed"][0][c],
                   res["fp32_mixed"][1][c]) for c in range(64)]
    keys = ("dE_rel_at_t_gate", "dQ_abs_at_t_gate", "dP_rel_at_t_gate", "max_dE_rel",
            "max_dQ_abs", "max_dP_rel")
    out["accuracy_fp32_mixed_vs_fp64"] = {
        k: (max(d[k] for d in per) if per[0][k] is not None else None) for k in keys}
    out["accuracy_fp32_mixed_vs_fp64"]["t_gate"] = per[0]["t_gate"]
    out["accuracy_fp32_mixed_vs_fp64"]["note"] = "maximum over the 64 configurations"
    return out


def ensemble_leg(qcbflow, args):
    """The 64-lattice ensemble workload (report_ensemble with its own defaults:
    150 steps, 30 sweeps, autotuned) condensed for the bench line: the
    FP32-mixed batched flow against the plain FP64 path flowing the same
    configurations one at a time, and the precision gates over all 64
    configurations."""
    ea = argparse.Namespace(**vars(args))
    ea.steps_given, ea.sweeps, ea.threads_per_site, ea.no_autotune = False, None, None, False
    r = report_ensemble(qcbflow, ea)
    acc = r["accuracy_fp32_mixed_vs_fp64"]
    return {
        "workload": r["config"]["workload"] + " (hipGraph-replayed steps, autotuned launch shape)",
        "sum_sites": r["config"]["sum_sites"],
        "fp64_value": r["fp64"]["value"],
        "fp32_mixed_value": r["fp32_mixed"]["value"],
        "fp64_plain_one_at_a_time_value": r["fp64"]["sequential_value"],
        "fp32_mixed_plain_one_at_a_time_value": r["fp32_mixed"]["sequential_value"],
        "fp32_mixed_batched_vs_fp64_plain": r["fp32_mixed_batched_vs_fp64_plain"],
        "fp32_mixed_vs_fp64_batched": r["fp32_mixed_vs_fp64"],
        "bit_identical_to_single_lattice_path": bool(
            r["fp64"]["bit_identical_to_single_lattice_path"]
            and r["fp32_mixed"]["bit_identical_to_single_lattice_path"]),
        "launch_shape": {"fp64": r["fp64"]["launch_shape"],
                         "fp32_mixed": r["fp32_mixed"]["launch_shape"]},
        "accuracy_fp32_mixed_vs_fp64": acc,
        "gate_dE_lt_1e-6": acc["dE_rel_at_t_gate"] is not None and acc["dE_rel_at_t_gate"] < 1e-6,
        "gate_dQ_lt_1e-4": acc["dQ_abs_at_t_gate"] is not None and acc["dQ_abs_at_t_gate"] < 1e-4,
    }


def headline(qcbflow, args):
    """The single-lattice workload of the bench line (a preset, by default the
    64^4 lattice measured every 10 steps): FP64 timed leg (value), FP32-mixed
    timed leg, accuracy leg to t = 2, flow-only and PCIe-inclusive rates, the
    copy kernel, and the CPU baseline on the same field."""
    Lp, beta_p, seed_p, sweeps_p, every_p, K_p = PRESETS[args.preset]
    L = lattice_of(args, Lp)
    V = volume(L)
    K = args.steps if args.steps_given else K_p
    beta = args.beta if args.beta is not None else beta_p
    seed = args.seed if args.seed is not None else seed_p
    sweeps = args.sweeps if args.sweeps is not None else sweeps_p
    every = every_p if K % every_p == 0 else math.gcd(K, every_p)
    lat = "x".join(map(str, L))
    log(f"lattice {L}, V = {V}; hot start seed {seed}, {sweeps} heatbath sweeps at beta {beta}; "
        f"{K} timed RK3 steps, measured every {every}")
    t0 = time.perf_counter()
    src = thermalised(qcbflow, L, beta, seed, sweeps)
    log(f"thermalised in {time.perf_counter() - t0:.2f} s, plaquette {src.plaquette():.6f}")

    gf = qcbflow.GradientFlow(L, args.eps)
    gf.copy_from(src)
    shape64 = setup_shape(qcbflow, gf, args)
    log(f"fp64 launch shape (map mode, window, tile code, fused, threads/site, lds): {shape64}")

    # timed region: K RK3 steps, K/every + 1 measurements, every launch timed
    secs, tqe, plaq, ms, launches = timed_flow(gf, src, K, args.warmup, every, args.time_every)
    value = V * K / secs
    log(f"timed: {K} steps in {secs:.4f} s -> {value:.4e} site-steps/s; "
        f"E(t={tqe[-1, 0]:.2f}) = {tqe[-1, 2]:.6g}, Q = {tqe[-1, 1]:.6g}")

    avg_stage_ms, achieved, bytes_per_launch = stage_roofline(ms, launches, V, "fp64")
    per_stage = {f"stage{k + 1}_avg_ms": ms[k] / launches[k] for k in range(3) if launches[k]}
    meas_avg_ms = ms[3] / max(1, launches[3])
    flops_rate = V * FLOW_FLOPS_PER_SITE_STEP / (3 * avg_stage_ms * 1e-3) / 1e12
    traffic = traffic_from_profiles(L)

    # flow-only throughput (no measurements)
    gf.synchronize()
    t2 = time.perf_counter()
    gf.rk3_step(K)
    gf.synchronize()
    flow_only = V * K / (time.perf_counter() - t2)

    # the same series when the field comes from and goes back to host memory
    gf.synchronize()
    t3 = time.perf_counter()
    host_links = src.field()
    t_down = time.perf_counter() - t3
    t3 = time.perf_counter()
    gf.copy(host_links)
    gf.synchronize()
    t_up = time.perf_counter() - t3
    pcie_value = V * K / (secs + t_up + t_down)
    log(f"host transfers of the field: up {t_up * 1e3:.1f} ms, down {t_down * 1e3:.1f} ms -> "
        f"{pcie_value:.4e} site-steps/s including them")

    # a plain device copy of 1 GiB (or the calibration size): the bandwidth a
    # simple streaming kernel reaches on this GPU, beside the published peak
    copy_mib = args.calibrate_copy_mib if args.calibrate_copy_mib > 0 else 1024
    copy_ms = gf.copy_bench(copy_mib << 20, 10)
    copy_gbs = 2 * (copy_mib << 20) / (copy_ms * 1e-3) / 1e9
    log(f"copy kernel: {copy_mib} MiB in {copy_ms:.4f} ms -> {copy_gbs:.1f} GB/s (read+write)")
    gf.close()

    # the FP32-mixed form on the same input and cadence
    mixed = None
    if not args.no_mixed:
        gm = qcbflow.GradientFlow(L, args.eps, qcbflow.FP32_MIXED)
        gm.copy_from(src)
        shape32 = setup_shape(qcbflow, gm, args)
        msecs, mtqe, mplaq, mms, mn = timed_flow(gm, src, K, args.warmup, every, args.time_every)
        mstage, mach, _ = stage_roofline(mms, mn, V, "fp32_mixed")
        mixed = {
            "value": V * K / msecs,
            "vs_fp64": (V * K / msecs) / value,
            "ms_per_step": msecs / K * 1e3,
            "stage_avg_ms": mstage,
            "measure_avg_ms": mms[3] / max(1, mn[3]),
            "stage1_measure": fused_stats(mms, mn, V, "fp32_mixed"),
            "hbm_achieved_gbs": mach,
            "hbm_frac": mach / HBM_PEAK_GBS,
            "bytes_per_site_step": MIXED_BYTES_PER_SITE_STEP,
            "fp32_tflops": V * MIXED_FLOPS_PER_SITE_STEP / (3 * mstage * 1e-3) / 1e12,
            "launch_shape": shape32,
            "exp_warnings": gm.exp_warnings(),
            "device_memory": device_memory(V, "fp32_mixed"),
        }
        mixed["fp32_frac"] = mixed["fp32_tflops"] / FP32_PEAK_TFLOPS
        mixed["fp32_frac_unpacked"] = mixed["fp32_tflops"] / FP32_UNPACKED_TFLOPS
        log(f"fp32-mixed: {mixed['value']:.4e} site-steps/s ({mixed['vs_fp64']:.2f}x fp64), "
            f"stage {mstage:.4f} ms")
        gm.close()
        # the precision gates at t = 2 (FP32-mixed vs FP64), on a leg of its own (2/eps steps)
        acc = accuracy_leg(qcbflow, src, L, args.eps, every)
        mixed["accuracy_vs_fp64"] = acc
        log(f"fp32-mixed vs fp64 at t={acc['t_gate']:.2f} ({acc['rk3_steps']} steps): "
            f"dE/E {acc['dE_rel_at_t_gate']:.2e}, dQ {acc['dQ_abs_at_t_gate']:.2e}, "
            f"dP/P {acc['dP_rel_at_t_gate']:.2e}")
    src.close()

    # the batched form on its own workload (64 lattices of uneven volumes)
    ensemble = None
    if not args.no_mixed and not args.no_ensemble and args.L is None and args.preset == "large":
        ensemble = ensemble_leg(qcbflow, args)

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        nproc, model = host_cpus()
        threads = min(args.cpu_threads, nproc)
        log(f"cpu baseline: {args.cpu_steps} RK3 step(s) + 1 measurement on {threads} of "
            f"{nproc} CPUs ({model}) ...")
        cpu = cpu_baseline(host_links, L, args.eps, every, threads, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.4e} {UNIT}")
    del host_links

    return {
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
        "data": f"synthetic: Haar-random SU(3) hot start (seed {seed}) thermalised by "
                f"{sweeps} heatbath sweeps at beta={beta}, generated on the GPU",
        "config": {
            "workload": f"Wilson flow FP64, {lat} lattice ({V} sites), eps={args.eps}, "
                        f"plaquette+Q+E measured every {every} RK3 step(s)",
            "lattice": list(L),
            "sites": V,
            "eps": args.eps,
            "beta": beta,
            "seed": seed,
            "heatbath_sweeps": sweeps,
            "measure_every_rk3_steps": every,
            "precision": "fp64",
            "parallelism": "single GPU",
            "launch_shape": shape64,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_rk3_stage_lds<" if shape64[5] else "k_rk3_stage<f64, ")
                      + "|".join(str(k + 1) for k in range(3) if launches[k])
                      + "> (one whole RK3 stage per launch"
                      + ("; the stage 1 after a measurement runs fused with it and is "
                         "reported as stage1_measure)" if launches[4] else ")"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_stage_ms,
            "timed_launches": launches[0] + launches[1] + launches[2],
            "timed_every_nth_launch": args.time_every,
            "fp64_tflops": flops_rate,
            "fp64_frac": flops_rate / FP64_PEAK_TFLOPS,
            "copy_kernel_gbs": copy_gbs,
            "traffic_gbs": None if traffic is None else traffic / (avg_stage_ms * 1e-3) / 1e9,
            **per_stage,
        },
        "measure_avg_ms": meas_avg_ms,
        "stage1_measure": fused_stats(ms, launches, V, "fp64"),
        "flow_only_value": flow_only,
        "pcie_inclusive_value": pcie_value,
        "host_transfer_ms": {"up": t_up * 1e3, "down": t_down * 1e3},
        "device_memory": device_memory(V, "fp64"),
        "fp32_mixed": mixed,
        "batched_ensemble": ensemble,
        "cpu_baseline": cpu,
        "copy_calibration": {"mib": copy_mib, "avg_ms": copy_ms},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed RK3 steps (default: 400 for the 64^4 workload, 200 for c2; "
                         "250/150 for the reports)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--preset", choices=sorted(PRESETS), default="large",
                    help="single-lattice workload: large = 64^4, beta 6.4, seed 2026, measured "
                         "every 10 steps (default); c2 = 32^4, beta 6.0, seed 7, every step")
    ap.add_argument("--L", type=int, nargs="+", default=None)
    ap.add_argument("--eps", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--beta", type=float, default=None)
    ap.add_argument("--sweeps", type=int, default=None,
                    help="heatbath sweeps of the input (default: the workload's)")
    ap.add_argument("--report", choices=("precision", "ensemble"), default=None)
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="oracle RK3 steps of the CPU baseline (about 12 s each at 64^4 on 16 "
                         "host threads)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-mixed legs")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-lattice ensemble leg of the default bench line")
    ap.add_argument("--time-every", type=int, default=1,
                    help="time every n-th launch of each kernel class with HIP events")
    ap.add_argument("--map-mode", type=int, default=None,
                    help="workgroup->tile order (qcbf_tune); default: the library's")
    ap.add_argument("--map-win", type=int, default=None)
    ap.add_argument("--tile-shape", type=int, default=None, help="qcbf_tune tile-shape code")
    ap.add_argument("--threads-per-site", type=int, choices=(1, 2, 4), default=None,
                    help="threads per site of the RK3 stage kernels (default 4)")
    ap.add_argument("--lds-stage", type=int, choices=(0, 1), default=None,
                    help="stage kernel: 1 = shared links staged in LDS (default), 0 = plain")
    ap.add_argument("--autotune", action="store_true",
                    help="single lattice: run qcbf_autotune before the warmup (default: the "
                         "library's launch shape)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="ensemble report: skip qcbf_batch_autotune")
    ap.add_argument("--no-fuse", action="store_true",
                    help="measure with a separate kernel instead of inside the next RK3 stage 1")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run the plain copy kernel on this many MiB (default 1024)")
    args = ap.parse_args()
    args.steps_given = args.steps is not None

    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (--gpus {args.gpus} requested); "
              "the whole lattice lives on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import qcbflow

    name, cus = qcbflow.device_info()
    log(f"device {name}, {cus} CUs")
    if args.report == "precision":
        out = report_precision(qcbflow, args)
    elif args.report == "ensemble":
        out = report_ensemble(qcbflow, args)
    else:
        out = headline(qcbflow, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
