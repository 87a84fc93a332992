#!/usr/bin/env python3
"""Throughput of the propagation path on one MI355X.

Headline workload (default): 2048 cosZ in [-1, 1] x 2048 energies log-spaced 0.05-1000 GeV on
the 12-shell jittered PREM-like Earth; one step = the four variants nu/nubar x normal/inverted
ordering, i.e. four calculateProbabilities of the full grid (16.8 M pairs) through the drop-in
API, each with the reference's per-call parameter hand-over (cudapropagator.cuh:191-220), inputs
already resident in HBM, results left in HBM (SoA, FP32 for the mixed kernel).

Prints one JSON line.
  value     pairs/s of the whole step sequence (host launch overhead included).
  roofline  the pair kernel: flops of the algorithm it runs (cudaprob3_amd/metrics.py,
            within ~1 % of rocprofv3's instruction counters) per launch / the average duration
            of one launch, from back-to-back launches between HIP events on its stream
            (cp3h_time_kernels); the reference algorithm's flop rate for the same results is
            `reference_work_TFLOPs`.
  scan1024, mcmc8192, surrogate
            the batched parameter scan, the Markov chain over the 8192^2 grid on 100 shells and
            the learned-surrogate comparison, each with its own roofline and CPU baseline.
  cpu_baseline
            the reference CpuPropagator<double> (or the oracle restatement if the reference
            build is absent) on the same four grids, with the host threads of the box's share;
            the same leg compares every GPU result with the reference's FP64 values
            (accuracy_vs_reference_fp64).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Threads for the CPU baselines: every CPU this process may run on (its affinity set),
    capped by OMP_NUM_THREADS when that is set (the CPU share the box grants one GPU)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(aff, int(env)) if env.isdigit() and int(env) > 0 else aff


def cpu_fields(threads: int) -> dict:
    return {"cores": threads, "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def reference_fn():
    import oracle
    kind = "reference" if oracle.ref_available() else "port"
    return kind, (oracle.ref_propagate if kind == "reference" else oracle.propagate)


def make_prop(g, precision):
    """A propagator holding one variant's grid, Earth model and hypothesis."""
    from cudaprob3_amd.propagator import HipPropagatorSingle
    p = HipPropagatorSingle(0, len(g.cosines), len(g.energies), precision=precision)
    p.setEnergyList(g.energies)
    p.setCosineList(g.cosines)
    p.setDensity(g.radii, g.rhos)
    p.setProductionHeight(g.h_km)
    p.setMNSMatrix(*g.params[:4])
    p.setNeutrinoMasses(*g.params[4:])
    return p


def time_steps(props, grids, steps, warmup, sync, after_warmup=None, schedule="rolling"):
    """One step = one calculateProbabilitiesAsync (parameter hand-over + launches on the
    propagator's stream) and one waitForCompletion of every variant.  schedule "grouped": all
    launched, then all awaited, before the next step; "rolling": each propagator is awaited and
    relaunched in turn, so it starts its next computation as soon as its previous one is done
    while the others still run (every computation is still awaited before its propagator is used
    again, and all are awaited at the end).  after_warmup() runs between the warm-up and the
    timed steps (untimed)."""
    def grouped_step():
        for p, g in zip(props, grids):
            p.calculateProbabilitiesAsync(g.antinu)
        for p in props:
            p.waitForCompletion()

    def rolling_step():
        for p, g in zip(props, grids):
            p.waitForCompletion()
            p.calculateProbabilitiesAsync(g.antinu)

    step = rolling_step if schedule == "rolling" else grouped_step
    for _ in range(warmup):
        step()
    for p in props:
        p.waitForCompletion()
    if after_warmup is not None:
        after_warmup()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    for p in props:
        p.waitForCompletion()
    sync()
    return time.perf_counter() - t0


def kernel_times(props, grids, reps, preroll_s=0.25):
    """Average duration of one set-up and one pair-kernel launch over the variants, each kernel
    launched `reps` times back to back between two HIP events (cp3h_time_kernels), after at
    least `preroll_s` seconds of untimed back-to-back launches (the clocks of a GPU that has been
    idle or lightly loaded take tens of milliseconds of steady work to come up)."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < preroll_s:
        for p, g in zip(props, grids):
            p.time_kernels(g.antinu, 10)
    setup, pairs = [], []
    for p, g in zip(props, grids):
        su, pa = p.time_kernels(g.antinu, reps)
        setup.append(su)
        pairs.append(pa)
    return float(np.mean(setup)), float(np.mean(pairs))


def d2h_times(prop, reps=5):
    """Copy-back of one grid's results into the handle's pinned host buffer
    (cp3h_fetch_results, the lazy copy of the first getProbability): best of `reps`."""
    from cudaprob3_amd import _lib
    lib = _lib.load()
    ptr, nbytes = ctypes.c_void_p(), ctypes.c_int()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.check(lib.cp3h_fetch_results(prop._handle(), ctypes.byref(ptr), ctypes.byref(nbytes)))
        best = min(best, time.perf_counter() - t0)
    size = 9 * prop.n_cosines * prop.n_energies * nbytes.value
    return best * 1e3, size


def accuracy_vs_fp64_kernel(grids, props):
    """max / rms |dP| of the mixed kernel against the FP64 kernel over the full grids (the FP64
    kernel is itself held within 1e-12 of the reference by the parity tests)."""
    worst, sq, n = 0.0, 0.0, 0
    for g, p_mixed in zip(grids, props):
        p64 = make_prop(g, "fp64")
        p64.calculateProbabilities(g.antinu)
        p_mixed.calculateProbabilities(g.antinu)
        d = np.abs(p64.getProbabilities() - p_mixed.getProbabilities().astype(np.float64))
        p64.close()
        worst = max(worst, float(d.max()))
        sq += float((d * d).sum())
        n += d.size
    return worst, float(np.sqrt(sq / n))


VALU_CEILING: dict = {}


def measure_valu_ceiling() -> dict:
    """The sustained non-packed FP32 / FP64 FMA rates of the GPU (cp3h_measure_valu_peak), measured
    once per run: the ceiling of kernels that issue one FMA per instruction, reported beside the
    data-sheet peak."""
    from cudaprob3_amd import _lib
    f32, f64 = ctypes.c_float(), ctypes.c_float()
    _lib.check(_lib.load().cp3h_measure_valu_peak(0, ctypes.byref(f32), ctypes.byref(f64)))
    VALU_CEILING.update({"fp32_fma_TFLOPs": f32.value, "fp64_fma_TFLOPs": f64.value})
    return VALU_CEILING


def with_valu_ceiling(roof: dict, arith: str = "fp32") -> dict:
    """Adds the measured FMA ceiling of `arith` ("fp32" or "fp64"), and the fraction of it the
    kernel reaches, to a roofline entry."""
    key = f"{arith}_fma_TFLOPs"
    if VALU_CEILING and roof.get("achieved"):
        roof["measured_fma_ceiling_TFLOPs"] = VALU_CEILING[key]
        roof["frac_of_measured_fma_ceiling"] = roof["achieved"] / VALU_CEILING[key]
    return roof


def traffic_from_profile(workload, precision, key="hbm_bytes_per_launch"):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def cpu_baseline_and_accuracy(grids, gpu_results, threads):
    """The reference CpuPropagator<double> on the same four grids (timed), and every GPU result
    of the mixed kernel (gpu_results, SoA per grid; None: timing only) against the reference's
    FP64 values."""
    kind, fn = reference_fn()
    dt, worst, sq, n = 0.0, 0.0, 0.0, 0
    for i, g in enumerate(grids):
        t0 = time.perf_counter()
        ref = fn(g.antinu, g.cosines, g.energies, g.radii, g.rhos, g.h_km, g.params, threads=threads)
        dt += time.perf_counter() - t0
        if gpu_results is None:
            continue
        d = np.abs(gpu_results[i].transpose(1, 2, 0).astype(np.float64) - ref)      # [ic, ie, t]
        worst = max(worst, float(d.max()))
        sq += float(np.einsum("ijk,ijk->", d, d))
        n += d.size
        del ref, d
    pairs = sum(g.n_pairs for g in grids)
    base = {"value": pairs / dt, "unit": "pairs/s", **cpu_fields(threads), "kind": kind,
            "sample": f"one full step: {len(grids)} grids of {len(grids[0].cosines)}x{len(grids[0].energies)} "
                      f"pairs, CpuPropagator<double> algorithm, OpenMP dynamic over cosines, {dt:.1f} s"}
    if gpu_results is None:
        return base, None
    acc = {"max_abs_dP": worst, "rms_dP": float(np.sqrt(sq / n)), "pairs": pairs, "channels": 9,
           "against": f"{kind} CpuPropagator<double>, every pair of the step"}
    return base, acc


def run_jitter12(args, sync):
    """The headline line: the four 2048^2 variants per step."""
    from cudaprob3_amd import metrics as M
    from cudaprob3_amd import workloads as W
    import torch

    grids = W.jitter12_variants(args.n)
    workload = f"jitter12_{args.n}"
    pairs_per_step = sum(g.n_pairs for g in grids)
    log(f"[bench] workload {workload}: {len(grids)} variants x {grids[0].n_pairs} pairs, precision {args.precision}")
    # each propagator enqueues on its own stream (created by the handle, default priority), as
    # each of the reference's CudaPropagatorSingle objects does, so one variant's set-up kernel
    # can run beside the previous variant's pair kernel.  The kernel durations of the roofline
    # are timed separately, one propagator at a time (kernel_times), where nothing overlaps.
    # --shared-stream puts all four on one stream instead.
    props = [make_prop(g, args.precision) for g in grids]
    shared = torch.cuda.Stream() if args.shared_stream else None
    if shared is not None:
        for p in props:
            p.set_stream(shared.cuda_stream)

    # kernel durations between the warm-up and the timed steps: back-to-back launches, which also
    # keep the clocks up for the timed steps
    times = {}
    dt = time_steps(props, grids, args.steps, args.warmup, sync,
                    after_warmup=lambda: times.update(k=kernel_times(props, grids, args.kernel_reps)))
    # the same steps with all four launched and all four awaited before the next step
    dt_grouped = time_steps(props, grids, args.steps, 0, sync, schedule="grouped")
    setup_ms, k_ms = times["k"]
    value = args.steps * pairs_per_step / dt
    value_grouped = args.steps * pairs_per_step / dt_grouped
    log(f"[bench] {args.steps} steps in {dt:.4f} s -> {value:.4e} pairs/s (each propagator relaunched as soon as "
        f"it is done); {value_grouped:.4e} pairs/s with all four awaited before the next step")

    ml = M.maxlayers_of(grids[0].cosines, grids[0].radii, grids[0].rhos)
    ne = len(grids[0].energies)
    elem = 4 if args.precision == "mixed" else 8
    out_bytes = elem * 9 * grids[0].n_pairs
    ref_flops = M.grid_flops(ml, ne)
    roof = {"bound": "valu"}
    if args.precision == "mixed":
        f32, f64 = M.mixed_grid_flops(ml, ne)
        roof.update(M.valu_roofline(f32, f64, k_ms))
        kernel = "pairs_mixed_kernel<false>"
    else:
        # FP64 kernels against the FP64 vector peak: the plain kernel runs the reference
        # algorithm itself, so its count is the reference's; the hoisted kernel's is its own
        if args.precision == "fp64_plain":
            flops, basis = ref_flops, "reference-algorithm flops (the kernel runs that algorithm)"
        else:
            flops, basis = float(M.f64_hoisted_flops_per_pair(ml).sum()) * ne, "the hoisted FP64 kernel's flops"
        a = flops / (k_ms * 1e-3) / 1e12
        roof.update({"achieved": a, "peak": M.PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": a / M.PEAK_FP64_VALU_TFLOPS, "basis": basis, "flops_per_launch": {"f64": flops}})
        kernel = "propagate_plain_f64" if args.precision == "fp64_plain" else "pairs_f64_kernel"
    roof.update({"traffic": traffic_from_profile(workload, args.precision), "kernel": kernel, "kernel_ms": k_ms,
                 "setup_kernel_ms": setup_ms, "launch_reps": args.kernel_reps,
                 "reference_work_TFLOPs": ref_flops / (k_ms * 1e-3) / 1e12,
                 "output_GBps": out_bytes / (k_ms * 1e-3) / 1e9})
    counted = traffic_from_profile(workload, args.precision, "counter_flops_per_launch")
    if counted:
        roof["counter_flops_per_launch"] = counted
    log(f"[bench] pair kernel {k_ms * 1e3:.1f} us, set-up {setup_ms * 1e3:.1f} us per launch "
        f"({args.kernel_reps} back to back); {roof['achieved']:.1f} TFLOP/s = {roof['frac']:.3f} of {roof['peak']}; "
        f"output {roof['output_GBps']:.0f} GB/s")

    d2h_ms, d2h_bytes = d2h_times(props[0])
    log(f"[bench] copy-back of one grid ({d2h_bytes / 1e6:.0f} MB) into pinned memory: {d2h_ms:.2f} ms")
    result = {
        "metric": "pairs/s: exact mixed-FP32 vs surrogate at max|dP|<=1e-3 vs FP64, 2048^2 grid",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 eigenvalues and phase reduction)" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{workload}: {args.n} cosZ in [-1,1] x {args.n} E log 0.05-1000 GeV, "
                               "12 jittered PREM shells, h=22 km, nu/nubar x NO/IO (4 grids per step)",
                   "precision": args.precision, "pairs_per_step": pairs_per_step,
                   "streams": "one shared" if args.shared_stream else "one per propagator",
                   "steps": "each propagator awaited and relaunched in turn (rolling); every computation "
                            "is awaited before its propagator is reused and all at the end"},
        "value_grouped_steps": value_grouped,
        "roofline": with_valu_ceiling(roof, "fp32" if args.precision == "mixed" else "fp64"),
        "d2h": {"ms_per_grid": d2h_ms, "bytes_per_grid": d2h_bytes, "GBps": d2h_bytes / (d2h_ms * 1e-3) / 1e9,
                "pairs_per_s_including_d2h": grids[0].n_pairs / ((dt / args.steps / len(grids)) + d2h_ms * 1e-3),
                "note": "results stay in HBM in the timed steps; this is the lazy copy-back of getProbability"},
    }
    gpu_results = None
    if args.precision == "mixed":
        if not args.no_accuracy:
            worst, rms = accuracy_vs_fp64_kernel(grids, props)
            result["accuracy_vs_fp64_kernel"] = {"max_abs_dP": worst, "rms_dP": rms}
            log(f"[bench] mixed vs fp64 kernel: max|dP| {worst:.3e} rms {rms:.3e}")
        if not args.no_cpu_baseline:
            gpu_results = []
            for p, g in zip(props, grids):
                p.calculateProbabilities(g.antinu)
                gpu_results.append(p.getProbabilities().copy())
    if not args.no_extra:
        extra, extra_roof = {}, {}
        for prec in ("fp64", "fp64_plain"):
            if prec == args.precision:
                continue
            qs = [make_prop(g, prec) for g in grids]
            dtq = time_steps(qs, grids, 2, 1, sync)
            extra[prec] = 2 * pairs_per_step / dtq
            # the FP64 pair kernels against the FP64 vector peak, on the flops of the algorithm each
            # runs (the plain kernel's is the reference's own)
            su, ka = kernel_times(qs, grids, 5, preroll_s=0.1)
            flops = ref_flops if prec == "fp64_plain" else float(M.f64_hoisted_flops_per_pair(ml).sum()) * ne
            a = flops / (ka * 1e-3) / 1e12
            extra_roof[prec] = with_valu_ceiling(
                {"kernel": "propagate_plain_f64" if prec == "fp64_plain" else "pairs_f64_kernel",
                 "kernel_ms": ka, "setup_kernel_ms": su, "flops_per_launch": flops, "achieved": a,
                 "peak": M.PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s", "frac": a / M.PEAK_FP64_VALU_TFLOPS}, "fp64")
            log(f"[bench] {prec}: {extra[prec]:.4e} pairs/s; pair kernel {ka:.3f} ms, {a:.1f} TFLOP/s = "
                f"{a / M.PEAK_FP64_VALU_TFLOPS:.3f} of the FP64 vector peak")
            for q in qs:
                q.close()
        result["other_precisions_pairs_per_s"] = extra
        result["other_precisions_roofline"] = extra_roof
    for p in props:
        p.close()
    del shared
    return result, grids, gpu_results


def run_scan(args, sync, steps, warmup, cpu=True):
    """Batched scan: 1024 sets (seeded), each step = new hypotheses for every set + one
    set-up launch + one persistent pair launch, replayed as a captured HIP graph."""
    from cudaprob3_amd import metrics as M
    from cudaprob3_amd import workloads as W
    from cudaprob3_amd.batch import BatchPropagator, SetSpec

    specs = [SetSpec(s.cosines, s.energies, s.radii, s.rhos, s.params, s.antinu, s.h_km) for s in W.scan_sets()]
    b = BatchPropagator(specs, graph=False)
    waves, gpc, tuned_ms = b.autotune()
    log(f"[bench] scan1024: {b.n_pairs} pairs, autotuned {waves} waves/workgroup x {gpc} workgroups/CU "
        f"({tuned_ms:.3f} ms pair kernel)")
    setup_ms, k_ms = b.time_kernels(args.kernel_reps_batch)
    b.use_graph(True)
    rng = np.random.Generator(np.random.PCG64(33))
    hyps = [np.array([W.random_params(rng) for _ in specs]) for _ in range(steps + warmup)]
    # pipelined steps: the hypotheses of step k+1 are written (second descriptor array) while
    # the device runs step k; every step still computes every set with its new hypothesis
    for k in range(warmup):
        b.set_params_many(hyps[k])
        b.compute_async()
    b.wait()
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        b.set_params_many(hyps[warmup + k])
        b.compute_async()
    b.wait()
    sync()
    dt = time.perf_counter() - t0
    value = steps * b.n_pairs / dt
    log(f"[bench] scan1024: {steps} steps in {dt:.4f} s -> {value:.4e} pairs/s; pair kernel {k_ms:.3f} ms, "
        f"set-up {setup_ms:.3f} ms per launch")
    f32 = f64 = ref = 0.0
    for h in b.hosts:
        a, c = M.mixed_grid_flops(h.maxlayers, h.n_energies, form="local")
        f32, f64 = f32 + a, f64 + c
        ref += M.grid_flops(h.maxlayers, h.n_energies)
    out_bytes = 36 * b.n_pairs
    roof = {"bound": "valu", **M.valu_roofline(f32, f64, k_ms),
            "traffic": traffic_from_profile("scan1024", "mixed"),
            "counter_flops_per_launch": traffic_from_profile("scan1024", "mixed", "counter_flops_per_launch"),
            "kernel": f"batch_pairs<{waves}, false>",
            "kernel_ms": k_ms, "setup_kernel_ms": setup_ms, "launch_reps": args.kernel_reps_batch,
            "reference_work_TFLOPs": ref / (k_ms * 1e-3) / 1e12, "output_GBps": out_bytes / (k_ms * 1e-3) / 1e9}
    result = {
        "metric": "pairs/s: batched parameter scan, mixed FP32/FP64, one persistent launch",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 eigenvalues and phase reduction)", "data": "synthetic",
        "config": {"workload": "scan1024: 1024 parameter sets, cosZ 64-512 x E 64-1024 each, 4-50 shell "
                               "random Earths, new hypotheses every step, HIP-graph replay, steps pipelined "
                               "(host writes step k+1 while the device runs step k)",
                   "pairs_per_step": b.n_pairs, "autotuned": {"waves_per_group": waves, "groups_per_cu": gpc}},
        "roofline": with_valu_ceiling(roof),
    }
    if cpu:
        threads = host_threads()
        kind, fn = reference_fn()
        t0 = time.perf_counter()
        pairs = 0
        for s in b.sets[:64]:
            fn(s.antinu, s.cosines, s.energies, s.radii, s.rhos, s.h_km, s.params, threads=threads)
            pairs += len(s.cosines) * len(s.energies)
        cdt = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": pairs / cdt, "unit": "pairs/s", **cpu_fields(threads), "kind": kind,
                                  "sample": f"the first 64 of the 1024 sets ({pairs} pairs), {cdt:.1f} s"}
        log(f"[bench] scan1024 CPU baseline {pairs / cdt:.4e} pairs/s ({threads} threads)")
    b.close()
    return result


def run_mcmc(args, sync, steps, warmup, binned_steps, cpu=True):
    """Chain of sequential proposals: each step re-propagates the 8192 x 8192 grid on the
    100-shell Earth (mixed precision) and reduces (P - D)^2 on the device."""
    from cudaprob3_amd import metrics as M
    from cudaprob3_amd import workloads as W
    from cudaprob3_amd.mcmc import Chain

    grid = W.fine100_grid(8192)
    chain = Chain(grid, W.NUFIT, sigma_p=0.1, seed=5)
    log(f"[bench] mcmc8192: {chain.n_pairs} pairs per step, 100 shells")
    chain.run(warmup)
    setup_ms, k_ms = chain.prop.time_kernels(grid.antinu, args.kernel_reps_batch, data_ptr=chain.data.data_ptr())
    chain.stats.step_ms.clear()
    sync()
    t0 = time.perf_counter()
    st = chain.run(steps)
    sync()
    dt = time.perf_counter() - t0
    value = steps * chain.n_pairs / dt
    p50, p99 = st.percentile(50), st.percentile(99)
    ml = M.maxlayers_of(grid.cosines, grid.radii, grid.rhos)
    f32, f64 = M.mixed_grid_flops(ml, len(grid.energies), likelihood=True)
    ref = M.grid_flops(ml, len(grid.energies))
    log(f"[bench] mcmc8192: {steps} steps in {dt:.3f} s -> {value:.4e} pairs/s; step p50 {p50:.3f} ms p99 {p99:.3f} ms; "
        f"pair kernel {k_ms:.3f} ms; accepted {st.accepted}/{st.steps}")
    result = {
        "metric": "pairs/s: Markov-chain steps re-evaluating the full grid (mixed FP32/FP64)",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": dt / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 eigenvalues and phase reduction)", "data": "synthetic",
        "config": {"workload": f"mcmc8192: {len(grid.cosines)} cosZ in [-1,1] x {len(grid.energies)} E log 0.1-1000 GeV, "
                               "100 equal PREM shells, Gaussian proposals (1% of range), device-side likelihood",
                   "pairs_per_step": chain.n_pairs},
        "latency_ms": {"p50": p50, "p99": p99, "mean": float(np.mean(st.step_ms))},
        "acceptance": st.accepted / max(st.steps, 1),
        "roofline": with_valu_ceiling({"bound": "valu", **M.valu_roofline(f32, f64, k_ms),
                     "traffic": traffic_from_profile("mcmc8192", "mixed"),
                     "counter_flops_per_launch": traffic_from_profile("mcmc8192", "mixed", "counter_flops_per_launch"),
                     "kernel": "pairs_mixed_kernel<true>",
                     "kernel_ms": k_ms, "setup_kernel_ms": setup_ms, "launch_reps": args.kernel_reps_batch,
                     "reference_work_TFLOPs": ref / (k_ms * 1e-3) / 1e12}),
    }
    chain.close()
    if cpu:
        # one Markov-chain step's propagation on a sample of the grid: every 16th cosine row
        # (512 rows from -1 to 1, so the deep 100-shell paths are in proportion) x all energies
        threads = host_threads()
        kind, fn = reference_fn()
        rows = grid.cosines[::16]
        t0 = time.perf_counter()
        fn(grid.antinu, rows, grid.energies, grid.radii, grid.rhos, grid.h_km, grid.params, threads=threads)
        cdt = time.perf_counter() - t0
        pairs = len(rows) * len(grid.energies)
        result["cpu_baseline"] = {"value": pairs / cdt, "unit": "pairs/s", **cpu_fields(threads), "kind": kind,
                                  "sample": f"every 16th cosine row x all {len(grid.energies)} energies "
                                            f"({pairs} pairs, 100 shells), {cdt:.1f} s"}
        log(f"[bench] mcmc8192 CPU baseline {pairs / cdt:.4e} pairs/s ({threads} threads)")
    if binned_steps > 0:
        result["binned_chains"] = binned_chain_comparison(binned_steps, sync, grid)
    return result


def binned_chain_comparison(binned_steps, sync, grid):
    """The same chain on bin-averaged probabilities (the grid's energies as bin centres): the
    exact kernel alone (16 Gauss-Legendre nodes per bin) against the surrogate network trained
    for the 100-shell Earth plus its exact fallback."""
    import torch
    from cudaprob3_amd import train_surrogate as T
    from cudaprob3_amd import workloads as W
    from cudaprob3_amd.mcmc import BinnedChain

    path = T.path_for("fine100")
    if not os.path.exists(path):
        return {"skipped": f"{os.path.basename(path)} not built (python -m cudaprob3_amd.train_surrogate --earth fine100)"}
    net, emap, meta = T.load(path, max_batch=8192)
    out, truth = {}, {}
    for method in ("exact", "surrogate"):
        ch = BinnedChain(grid, W.NUFIT, method, net=net, emap=emap, earth="fine100", sigma_p=0.1, seed=5)
        truth[method] = ch.data.clone()
        ch.run(2)
        ch.stats = type(ch.stats)()
        sync()
        t0 = time.perf_counter()
        st = ch.run(binned_steps)
        sync()
        dt = time.perf_counter() - t0
        out[method] = {"steps": st.steps, "p50_ms": st.percentile(50), "p99_ms": st.percentile(99),
                       "bins_per_s": st.steps * ch.n_bins / dt, "acceptance": st.accepted / max(st.steps, 1)}
        if method == "surrogate":
            out[method]["exact_fraction"] = ch.exact_fraction
        log(f"[bench] binned chain ({method}): p50 {out[method]['p50_ms']:.2f} ms, p99 {out[method]['p99_ms']:.2f} ms, "
            f"{out[method]['bins_per_s']:.4e} bins/s")
        ch.close()
    out["max_abs_dP_surrogate_vs_exact_at_truth"] = float((truth["surrogate"] - truth["exact"]).abs().max())
    out["network"] = meta
    log(f"[bench] binned chains: surrogate+fallback vs exact at the true parameters max|dP| "
        f"{out['max_abs_dP_surrogate_vs_exact_at_truth']:.3e}")
    net.close()
    return out


def surrogate_cpu_baseline(args):
    """Bin averages by the reference CpuPropagator<double>: one hypothesis, every 16th cosine row
    x all bins, 16 Gauss-Legendre nodes each (the exact method's work on the CPU)."""
    from cudaprob3_amd import surrogate as S
    from cudaprob3_amd import workloads as W
    threads = host_threads()
    kind, fn = reference_fn()
    cos = np.linspace(-1, 1, args.n)[::16]
    E, _ = S.gl_energies(np.geomspace(0.05, 1000, args.n))
    radii, rhos = W.jitter12_model()
    t0 = time.perf_counter()
    fn(0, cos, E, radii, rhos, W.PRODUCTION_HEIGHT_KM, W.NUFIT, threads=threads)
    dt = time.perf_counter() - t0
    bins = len(cos) * args.n
    log(f"[bench] CPU baseline {bins / dt:.4e} bins/s")
    return {"value": bins / dt, "unit": "bins/s", **cpu_fields(threads), "kind": kind,
            "sample": f"one hypothesis, every 16th cosine row x {args.n} bins x 16 nodes "
                      f"({len(cos) * len(E)} pairs), {dt:.1f} s"}


def surrogate_comparison(args, sync):
    """Binned probabilities of many hypotheses on the 2048^2 grid: the trained network with the
    exact fallback (bins whose error-map bound exceeds 1e-3 / 1.5), against the exact
    mixed-precision kernel alone (16 Gauss-Legendre nodes per bin, binned in registers) and the
    plain FP64 kernel; max |dP| of the combined result against the exact binned values on
    `surrogate_validate` hypotheses and against FP64 on one."""
    import ctypes
    import torch
    from cudaprob3_amd import _lib
    from cudaprob3_amd import surrogate as S
    from cudaprob3_amd import train_surrogate as T
    from cudaprob3_amd import workloads as W
    from cudaprob3_amd.propagator import HipPropagatorSingle

    n, n_hyp, slots = args.n, args.surrogate_hypotheses, 8
    net, emap, meta = T.load(max_batch=8192)
    cos = np.linspace(-1.0, 1.0, n)
    centers = np.geomspace(0.05, 1000.0, n)
    rng = np.random.Generator(np.random.PCG64(44))
    hyps = [W.random_params(rng) for _ in range(n_hyp)]
    nus = [int(rng.random() < 0.5) for _ in range(n_hyp)]
    # the network kernel alone (grid mode, no skipped bins), timed with HIP events on its stream
    from cudaprob3_amd import metrics as M
    dev = torch.device("cuda:0")
    cos_t = torch.as_tensor(cos.astype(np.float32), device=dev)
    l10_t = torch.as_tensor(np.log10(centers).astype(np.float32), device=dev)
    hyp_t = torch.tensor([list(hyps[j]) + [nus[j]] for j in range(slots)], dtype=torch.float32, device=dev)
    y_t = torch.empty((slots, 9, n, n), dtype=torch.float32, device=dev)
    net.forward_grid(cos_t, l10_t, hyp_t, None, y_t)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    sync()
    e0.record(net.stream)
    for _ in range(reps):
        net.forward_grid(cos_t, l10_t, hyp_t, None, y_t)
    e1.record(net.stream)
    sync()
    mlp_ms = e0.elapsed_time(e1) / reps
    rows = slots * n * n
    mlp_tflops = rows * M.MLP_FLOPS_PER_ROW / (mlp_ms * 1e-3) / 1e12
    mlp_exec_tflops = rows * M.MLP_EXECUTED_FLOPS_PER_ROW / (mlp_ms * 1e-3) / 1e12
    del y_t
    log(f"[bench] network kernel: {rows / (mlp_ms * 1e-3):.4e} rows/s, {mlp_tflops:.0f} TFLOP/s algorithmic "
        f"({mlp_tflops / M.PEAK_BF16_MFMA_TFLOPS:.3f} of the BF16 MFMA peak), {mlp_exec_tflops:.0f} executed "
        f"(two-term products)")

    ev = S.GridEvaluator(net, emap, cos, centers, tol=1e-3, safety=1.5, slots=slots)
    ex = S.ExactBinnedGrid(torch, cos, centers, slots=slots)
    bins = n_hyp * n * n
    ev.run(hyps[:slots], nus[:slots])                     # warm-up (code, graphs)
    ex.run(hyps[:slots], nus[:slots])
    sync()
    t0 = time.perf_counter()
    stats = ev.run(hyps, nus)
    sync()
    t_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ex.run(hyps, nus)
    sync()
    t_x = time.perf_counter() - t0
    log(f"[bench] surrogate+fallback {bins / t_s:.4e} bins/s ({stats['exact_bins'] / bins:.3f} of bins exact), "
        f"exact mixed binned {bins / t_x:.4e} bins/s")

    # accuracy: combined result vs the exact binned kernel (fallback bins are identical)
    n_val = min(args.surrogate_validate, n_hyp)
    worst, sq, cnt = 0.0, 0.0, 0
    for c0 in range(0, n_val, slots):
        idx = list(range(c0, min(c0 + slots, n_val)))
        keep = {}
        # the evaluator hands results over per class, in its own order: key them by hypothesis
        ev.run([hyps[j] for j in idx], [nus[j] for j in idx],
               consume=lambda sub, out: keep.update({sub[m]: out[m].clone() for m in range(len(sub))}))

        def cmp(sub, out):
            nonlocal worst, sq, cnt
            for m, jj in enumerate(sub):
                d = (keep.pop(jj) - out[m]).abs()
                worst = max(worst, float(d.max()))
                sq += float((d.double() ** 2).sum())
                cnt += d.numel()
        ex.run([hyps[j] for j in idx], [nus[j] for j in idx], consume=cmp)
    log(f"[bench] surrogate+fallback vs exact binned on {n_val} hypotheses: max|dP| {worst:.3e}")

    # one hypothesis in FP64 (hoisted kernel, unbinned, averaged here) and the plain FP64 rate
    E_nodes, w = S.gl_energies(centers)
    w64 = torch.as_tensor(w.astype(np.float64), device="cuda")
    j = 0
    radii, rhos = W.jitter12_model()
    res = {}
    for prec in ("fp64", "fp64_plain"):
        p = HipPropagatorSingle(0, n, len(E_nodes), precision=prec)
        p.setEnergyList(E_nodes)
        p.setCosineList(cos)
        p.setDensity(radii, rhos)
        p.setProductionHeight(22.0)
        p.setMNSMatrix(*hyps[j][:4])
        p.setNeutrinoMasses(*hyps[j][4:])
        p.calculateProbabilities(nus[j])
        t0 = time.perf_counter()
        p.calculateProbabilities(nus[j])
        res[prec] = time.perf_counter() - t0
        if prec == "fp64":
            full = torch.empty((9, n, n, 16), dtype=torch.float64, device="cuda")
            sync()
            _lib.check(p._lib.cp3h_copy_results_device(p._handle(), ctypes.c_void_p(full.data_ptr()), full.numel()))
            ref64 = (full * w64).sum(dim=-1)        # elementwise: torch's batched matmul over this
            #                                           4.8 GB float64 operand returned wrong values
            del full
        p.close()
    keep = {}
    ev.run([hyps[j]], [nus[j]], consume=lambda i, out: keep.update(s=out[0].double().clone()))
    ex.run([hyps[j]], [nus[j]], consume=lambda i, out: keep.update(x=out[0].double().clone()))
    d64_s = float((keep["s"] - ref64).abs().max())
    d64_x = float((keep["x"] - ref64).abs().max())
    del ref64, keep
    log(f"[bench] hypothesis 0 vs FP64 binned: surrogate+fallback {d64_s:.3e}, exact mixed {d64_x:.3e}; "
        f"fp64 {n * n / res['fp64']:.4e} bins/s, fp64_plain {n * n / res['fp64_plain']:.4e} bins/s")
    classes = {k: ev.fallback_fraction(k) for k in range(S.N_CLASSES)}
    ev.close()
    ex.close()
    net.close()
    return {
        "workload": f"{n_hyp} hypotheses (scan ranges, either ordering, nu/nubar) x {n} cosZ x {n} energy bins "
                    "(+-2.5 %, 16-point Gauss-Legendre), jittered 12-shell Earth",
        "bins_per_s": bins / t_s,
        "exact_fraction": stats["exact_bins"] / bins,
        "exact_fraction_by_class": classes,
        "max_abs_dP_vs_exact_binned": worst, "rms_dP_vs_exact_binned": (sq / max(cnt, 1)) ** 0.5,
        "validated_hypotheses": n_val,
        "max_abs_dP_vs_fp64_hyp0": d64_s,
        "exact_mixed_binned_bins_per_s": bins / t_x,
        "exact_mixed_binned_max_abs_dP_vs_fp64_hyp0": d64_x,
        "exact_fp64_bins_per_s": n * n / res["fp64"],
        "exact_fp64_plain_bins_per_s": n * n / res["fp64_plain"],
        "speedup_vs_exact_mixed": t_x / t_s,
        "speedup_vs_exact_fp64_plain": (n * n / res["fp64_plain"]) ** -1 * bins / t_s,
        "network": meta,
        "network_roofline": {"bound": "mfma", "kernel": "mlp_forward_fused (grid mode)",
                             "achieved": mlp_tflops, "peak": M.PEAK_BF16_MFMA_TFLOPS, "unit": "TFLOP/s",
                             "frac": mlp_tflops / M.PEAK_BF16_MFMA_TFLOPS, "kernel_ms": mlp_ms,
                             "rows_per_launch": rows, "algorithmic_flops_per_row": M.MLP_FLOPS_PER_ROW,
                             "executed_flops_per_row": M.MLP_EXECUTED_FLOPS_PER_ROW,
                             "executed_TFLOPs": mlp_exec_tflops,
                             "rows_per_s": rows / (mlp_ms * 1e-3)},
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="jitter12", choices=["jitter12", "scan1024", "mcmc8192", "surrogate2048"],
                    help="jitter12: the 2048^2 grid, four variants per step, with the scan1024, mcmc8192 and "
                         "surrogate lines as sub-objects; scan1024: 1024 parameter sets with uneven grids and "
                         "Earth models in one persistent launch per step; mcmc8192: Markov-chain steps over an "
                         "8192^2 grid on a 100-shell Earth; surrogate2048: the surrogate comparison alone")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64", "fp64_plain"])
    ap.add_argument("--n", type=int, default=2048, help="grid edge of the jitter12 and surrogate workloads")
    ap.add_argument("--kernel-reps", type=int, default=50,
                    help="back-to-back launches per variant for the kernel durations of the jitter12 line")
    ap.add_argument("--kernel-reps-batch", type=int, default=5,
                    help="back-to-back launches for the kernel durations of the scan1024 and mcmc8192 lines")
    ap.add_argument("--sub-steps", type=int, default=20, help="timed steps of the scan1024 sub-object")
    ap.add_argument("--chain-steps", type=int, default=10000,
                    help="timed steps of the mcmc8192 sub-object (the workload's chain of 10,000 proposals)")
    ap.add_argument("--sub-warmup", type=int, default=3)
    ap.add_argument("--shared-stream", action="store_true",
                    help="jitter12: run the four propagators on one stream (default: each on its own)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other precisions' throughput")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate-vs-exact comparison")
    ap.add_argument("--no-scan", action="store_true", help="jitter12: skip the scan1024 sub-object")
    ap.add_argument("--no-mcmc", action="store_true", help="jitter12: skip the mcmc8192 sub-object")
    ap.add_argument("--surrogate-hypotheses", type=int, default=256)
    ap.add_argument("--binned-chain-steps", type=int, default=10,
                    help="mcmc8192: steps of the binned exact and surrogate chains (0: skip)")
    ap.add_argument("--surrogate-validate", type=int, default=64,
                    help="hypotheses whose every bin is checked against the exact binned kernel")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one elementwise copy of this many MiB (reads and writes that many "
                         "bytes), the known-size kernel that scales HBM counters read by a profiler")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU "
              f"path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU (cuda:0)", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    sync = torch.cuda.synchronize
    t_start = time.perf_counter()
    if args.copy_calibration_mib:
        src = torch.ones(args.copy_calibration_mib * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        sync()
        torch.neg(src, out=dst)
        sync()
        log(f"[bench] calibration copy: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written")
        del src, dst
    try:
        ceiling = measure_valu_ceiling()
        log(f"[bench] measured FMA ceiling (non-packed, 8 wavefronts per SIMD): FP32 {ceiling['fp32_fma_TFLOPs']:.1f}, "
            f"FP64 {ceiling['fp64_fma_TFLOPs']:.1f} TFLOP/s")
    except RuntimeError as e:           # the roofline entries then carry the data-sheet peak only
        log(f"[bench] FMA ceiling not measured: {e}")

    if args.workload == "scan1024":
        print(json.dumps(run_scan(args, sync, args.steps, args.warmup, cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.workload == "mcmc8192":
        print(json.dumps(run_mcmc(args, sync, args.steps, args.warmup, args.binned_chain_steps,
                                  cpu=not args.no_cpu_baseline)), flush=True)
        return
    if args.workload == "surrogate2048":
        sur = surrogate_comparison(args, sync)
        bins = args.surrogate_hypotheses * args.n * args.n
        print(json.dumps({
            "metric": "bins/s: surrogate + exact fallback at max|dP|<=1e-3 (binned probabilities, 2048^2 grid)",
            "value": sur["bins_per_s"], "unit": "bins/s", "n_gpus": 1, "steps": args.surrogate_hypotheses,
            "warmup": 8, "ms_per_step": bins / sur["bins_per_s"] / args.surrogate_hypotheses * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 MFMA (fp32 accumulation) + f32/f64 exact fallback", "data": "synthetic",
            "config": {"workload": "surrogate2048: " + sur["workload"]},
            "roofline": sur["network_roofline"], "surrogate": sur,
            "cpu_baseline": None if args.no_cpu_baseline else surrogate_cpu_baseline(args),
            "wall_s": time.perf_counter() - t_start}), flush=True)
        return

    result, grids, gpu_results = run_jitter12(args, sync)

    def sub_object(key, fn):
        # a failing sub-object is reported in its place (and on stderr) instead of taking the
        # headline line with it
        try:
            result[key] = fn()
        except Exception as e:          # noqa: BLE001 -- recorded, not swallowed
            import traceback
            traceback.print_exc()
            result[key] = {"error": f"{type(e).__name__}: {e}"}

    if args.precision == "mixed":
        if not args.no_surrogate:
            sub_object("surrogate", lambda: surrogate_comparison(args, sync))
        if not args.no_scan:
            sub_object("scan1024", lambda: run_scan(args, sync, args.sub_steps, args.sub_warmup,
                                                    cpu=not args.no_cpu_baseline))
        if not args.no_mcmc:
            sub_object("mcmc8192", lambda: run_mcmc(args, sync, args.chain_steps, args.sub_warmup,
                                                    args.binned_chain_steps, cpu=not args.no_cpu_baseline))
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"[bench] CPU baseline with {threads} threads ...")
        base, acc = cpu_baseline_and_accuracy(grids, gpu_results, threads)
        if acc is not None:
            result["accuracy_vs_reference_fp64"] = acc
            log(f"[bench] mixed vs the reference FP64 CPU path, every pair: max|dP| {acc['max_abs_dP']:.3e} "
                f"rms {acc['rms_dP']:.3e}")
        result["cpu_baseline"] = base
        log(f"[bench] CPU baseline {base['value']:.4e} pairs/s")
    result["wall_s"] = time.perf_counter() - t_start
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
