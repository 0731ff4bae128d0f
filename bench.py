"""Benchmark of the MI355X MPGOS solver path.

Default workload (one GPU): the Lorenz ensemble, 262144 instances,
rho ~ U[0, 21], x0 = (10, 10, 10) + N(0, 0.01), adaptive RKCK45 at
rtol = atol = 1e-10 over t in [0, 10], FP64.  One "step" is one Solve() over
the whole batch with the inputs already resident in HBM: a device-to-device
reset of the initial state followed by the integration kernel.  The metric is
accepted instance-steps per second (one accepted Runge-Kutta step of one
instance).

Other workloads (--workload), each printed as the same JSON line:
  duffing_rk4     16384 forced Duffing oscillators, fixed-step RK4,
                  dt = 2pi/1000 over 64 forcing periods (64001 steps each);
  duffing_uneven  131072 Duffing oscillators with forcing frequency
                  eps ~ 10^U[-3,0] and horizons of 16..4096 periods, RKCK45 at
                  1e-9, event x' = 0 with up to 256 logged hits; the longest
                  instances take ~6e7 steps, so the default run is
                  host-sliced (--slice-steps per launch, bit-identical to one
                  Solve, progress after each slice); --form single | relaunch
                  | sorted time one Solve of the reference's one-launch form,
                  the relaunch form or the sorted form, with progress read
                  from the device while it runs;
  duffing_adjoint 65536 Duffing oscillators with a quadrature state, RK4 at
                  dt = 2pi/500 (or --algorithm rkck45 at 1e-9): forward Solve
                  and discrete adjoint (record + reverse), reporting the
                  adjoint/forward cost ratio;
  ring            4096 per-block rings of 256 coupled Duffing units with a
                  dense coupling matrix, RKCK45 at 1e-9 over 4 periods; the
                  coupling product on FP64 MFMA (metric: accepted
                  system-steps/s).

The default run (Lorenz) also carries the rest of the metric: "adjoint" (the
discrete adjoint of the Duffing ensemble, RK4 and RKCK45, record + reverse
over forward with the autotuned checkpoint interval; the record pass is itself
the forward Solve with checkpoints, so "gradient_run_extra_over_forward" is
what a gradient adds to the forward Solve it replaces) and "ring" (the per-block
ring on the MFMA coupling kernel with its CPU sample, parity and adjoint).

Printed (one JSON line): the contract fields plus
  roofline      -- algorithmic FLOP/s of the dominant kernel (FLOPs per
                   attempted step from the oracle's op count, or the stated
                   per-step count for the ring, x attempted steps, divided by
                   the kernel time measured with HIP events on the solver's
                   stream) against the MI355X FP64 peak (78.6 TFLOP/s for both
                   the vector ALUs and the matrix cores);
  cpu_baseline  -- the FP64 OpenMP oracle on this host's cores, on the full
                   workload or a stated bounded sample of it;
  parity        -- the GPU result compared bit for bit with that oracle run.
--copy-calibration MiB runs plain device-to-device copies (for scaling HBM
counters) and prints the bytes moved.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "massively-parallel-gpu-ode-solver_amd"))

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
ENGINE_NOTE = "FP64 vector ALUs, no matrix instructions (MI355X FP64 vector peak 78.6 TFLOP/s)"
ENGINE_NOTE_F32 = ("FP32 vector ALUs for the stages (MI355X FP32 vector peak 157.3 TFLOP/s); error norm, controller and "
                   "event control in FP64")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def threads_here():
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_oracle(run, min_seconds=2.0):
    """Runs the oracle (run() -> output dict with 'seconds' and 'nacc') until at
    least min_seconds of CPU time are measured, so short samples are not
    dominated by timer and thread start-up noise.  Returns the first run's
    outputs, the accepted steps and the seconds summed over all runs, and the
    number of runs."""
    first = run()
    steps, secs, runs = float(first["nacc"].sum()), float(first["seconds"]), 1
    while secs < min_seconds:
        o = run()
        steps += float(o["nacc"].sum())
        secs += float(o["seconds"])
        runs += 1
    return first, steps, secs, runs


def cpu_record(value, unit, sample):
    return {"value": value, "unit": unit, "cores": threads_here(), "kind": "port", "sample": sample,
            "cpu_model": cpu_model()}


def bitwise(a, b, keys):
    return all(np.array_equal(np.ascontiguousarray(a[k]).view(np.uint8), np.ascontiguousarray(b[k]).view(np.uint8)) for k in keys)


# --------------------------------------------------------------------------- per-thread workloads
def sliced_solve(s, budget, label):
    """Solve in slices of `budget` attempted steps per instance (bit-identical
    to one Solve); returns the summed kernel time in ms.  Each slice reports
    the instances still unfinished."""
    total = 0.0
    k = 0
    t0 = time.perf_counter()
    while True:
        left = s.solve_slice(budget)
        total += s.last_solve_ms()
        k += 1
        log(f"{label}: slice {k}: {left} instances unfinished, kernel {total / 1e3:.1f} s, wall {time.perf_counter() - t0:.1f} s")
        if left == 0:
            return total


def watched_solve(s, label, every=20.0):
    """One Solve() that may run for minutes, with progress from the work
    itself: each instance writes its accepted-step count when it finishes, so
    while the Solve runs a second stream copies those counts (cleared before
    the Solve) and the number of finished instances is reported.  Returns the
    Solve's kernel time in ms."""
    from mpgos.enums import ListOfVariables as V
    from mpgos.tuning import hip
    h = hip()
    h.hipStreamQuery.argtypes = [ctypes.c_void_p]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipMemsetAsync.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
    ptr, nbytes = s.device_ptr(V.AcceptedSteps)
    counts = np.empty(nbytes // 8, np.int64)
    side = ctypes.c_void_p()
    if h.hipStreamCreate(ctypes.byref(side)) != 0 or h.hipMemsetAsync(ptr, 0, nbytes, s.stream()) != 0:
        raise RuntimeError("watched_solve: stream or memset failed")
    try:
        s.Solve()
        t0 = last = time.perf_counter()
        while h.hipStreamQuery(s.stream()) != 0:
            time.sleep(0.5)
            if time.perf_counter() - last >= every:
                last = time.perf_counter()
                if h.hipMemcpyAsync(counts.ctypes.data, ptr, nbytes, 2, side) != 0 or h.hipStreamSynchronize(side) != 0:
                    raise RuntimeError("watched_solve: progress copy failed")
                log(f"{label}: {np.count_nonzero(counts)} of {counts.size} instances finished after {last - t0:.0f} s")
        if h.hipStreamSynchronize(s.stream()) != 0:
            raise RuntimeError("watched_solve: Solve failed")
    finally:
        h.hipStreamDestroy(side)
    return s.last_solve_ms()


def time_form(mpgos, spec, cfg, steps, warmup, precision=0, slice_steps=0, watch=False):
    """K timed steps (device-side input reset + Solve) of one launch configuration."""
    from mpgos import workloads as W
    from mpgos.tuning import DeviceSnapshot, apply
    s = mpgos.ProblemSolver(spec["system"], spec["nt"], spec["algorithm"], precision, ndo=spec.get("ndo", 0))
    try:
        W.load(s, spec)
        apply(s, cfg)
        snap = DeviceSnapshot(s)
        for _ in range(warmup if not watch else 0):
            snap.restore()
            if slice_steps:
                snap.sync()
                sliced_solve(s, slice_steps, f"warmup {cfg}")
            else:
                s.Solve()
        snap.sync()
        kernel_ms = []
        t0 = time.perf_counter()
        for i in range(steps):
            snap.restore()
            if slice_steps:
                snap.sync()
                kernel_ms.append(sliced_solve(s, slice_steps, f"{cfg}"))
                continue
            if watch:
                snap.sync()
                kernel_ms.append(watched_solve(s, f"{cfg}"))
                continue
            s.Solve()
            snap.sync()
            kernel_ms.append(s.last_solve_ms())
            if steps > 1 and kernel_ms[-1] > 2000:
                log(f"step {i}: kernel {kernel_ms[-1]:.1f} ms")
        t1 = time.perf_counter()
        snap.close()
        s.SynchroniseFromDeviceToHost(mpgos.ListOfVariables.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        out = W.read(s)
    finally:
        s.close()
    return {"wall_ms_per_step": (t1 - t0) * 1e3 / steps, "kernel_ms": float(np.mean(kernel_ms)),
            "kernel_ms_min": float(np.min(kernel_ms)), "out": out, "config": cfg}


def tune(mpgos, spec, precision=0):
    from mpgos import workloads as W
    from mpgos.tuning import autotune
    s = mpgos.ProblemSolver(spec["system"], spec["nt"], spec["algorithm"], precision, ndo=spec.get("ndo", 0))
    try:
        W.load(s, spec)
        best, _ = autotune(s, log=log)
    finally:
        s.close()
    return best


def flops_per_attempt(pyoracle, W, spec, sample_idx):
    sample = W.subset(spec, sample_idx)
    fl = pyoracle.run(spec["system"], sample, count_flops=True)
    return fl["flops"] / float(fl["nacc"].sum() + fl["nrej"].sum())


def per_thread(args, mpgos, W, pyoracle):
    prec = {"f64": 0, "f32": 1, "f32k": 2}[args.precision]
    if args.workload == "lorenz":
        spec = W.config1_lorenz(nt=args.nt or 262144)
        name = f"lorenz_rkck45_{spec['nt']}"
        cpu_idx = None                                  # full workload on the CPU
        desc = {"system": "lorenz", "algorithm": "RKCK45", "rtol": 1e-10, "atol": 1e-10, "t_end": 10.0}
    elif args.workload == "duffing_rk4":
        spec = W.config0_duffing_rk4(nt=args.nt or 16384)
        name = f"duffing_rk4_{spec['nt']}"
        cpu_idx = None                                  # full workload on the CPU (about 1 s on 16 cores)
        desc = {"system": "duffing", "algorithm": "RK4", "dt": "2pi/1000", "periods": 64}
    else:
        spec = W.config2_duffing_uneven(nt=args.nt or 131072)
        name = f"duffing_uneven_rkck45_{spec['nt']}"
        cpu_idx = np.arange(0, spec["nt"], 512)         # stratified: every 512th instance
        desc = {"system": "duffing_eps", "algorithm": "RKCK45", "rtol": 1e-9, "atol": 1e-9, "events": "x'=0, 256 logged"}
    forms = {}
    sl = args.slice_steps if args.workload == "duffing_uneven" else 0
    if sl and args.form == "single":        # the reference's form: one plain launch for the whole batch
        forms["single"] = time_form(mpgos, spec, {"mode": 0, "block": 64}, args.steps, 0, prec, 0, watch=True)
    if sl and args.form == "relaunch":      # device-chained relaunches, budget --slice-steps per launch
        rcfg = {"mode": 2, "block": 64, "relaunch": sl, "launches": 64}
        forms["relaunch"] = time_form(mpgos, spec, rcfg, args.steps, 0, prec, 0, watch=True)
    if sl and args.form == "sorted":        # pilot launch, device sort by estimated remaining steps, one final launch,
        scfg = {"mode": 3, "block": 64, "relaunch": 2000, "graph": 1}   # all captured in one hipGraph
        forms["sorted"] = time_form(mpgos, spec, scfg, args.steps, 0, prec, 0, watch=True)
    if args.form in ("auto", "plain"):
        forms["plain"] = time_form(mpgos, spec, {"mode": 0, "block": 64}, args.steps, args.warmup, prec, sl)
    if args.form in ("auto", "persistent"):
        pcfg = {"mode": 1, "block": 256, "round": 32}
        if sl and args.solo:       # longest interval first, the longest instances alone in one-wave workgroups
            pcfg = {"mode": 1, "block": 64, "round": 256, "waves_per_simd": 1, "order": 1, "solo": args.solo}
        forms["persistent"] = time_form(mpgos, spec, pcfg, args.steps, args.warmup, prec, sl)
    if args.form in ("auto", "relaunch") and not sl:
        forms["relaunch"] = time_form(mpgos, spec, {"mode": 2, "block": 256, "variant": 2, "relaunch": 200, "launches": 32},
                                      args.steps, args.warmup, prec)
    if args.form in ("auto", "sorted") and not sl:
        forms["sorted"] = time_form(mpgos, spec, {"mode": 3, "block": 64, "variant": 0, "relaunch": 400},
                                    args.steps, args.warmup, prec)
    if args.form == "auto" and not args.no_tune and not sl:
        best_cfg = tune(mpgos, spec, prec)
        forms["autotuned"] = time_form(mpgos, spec, best_cfg, args.steps, args.warmup, prec)
    best = min(forms, key=lambda k: forms[k]["wall_ms_per_step"])
    res = forms[best]
    out = res["out"]
    accepted = int(out["nacc"].sum())
    attempted = accepted + int(out["nrej"].sum())
    value = accepted / (res["wall_ms_per_step"] * 1e-3)

    nt = spec["nt"]
    if args.workload == "duffing_uneven":      # FLOPs per attempted step barely depend on the instance
        sample = np.argsort(spec["TD"][1])[:16]
    else:
        sample = np.arange(0, nt, max(1, nt // (16 if args.workload != "lorenz" else 1024)))
    fpa = flops_per_attempt(pyoracle, W, spec, sample)
    achieved = fpa * attempted / (res["kernel_ms"] * 1e-3) / 1e12

    cpu = parity = None
    if not args.no_cpu:
        sub = spec if cpu_idx is None else W.subset(spec, cpu_idx)
        ref, steps, secs, runs = timed_oracle(lambda: pyoracle.run(spec["system"], sub, nthreads=threads_here()))
        what = "full workload" if cpu_idx is None else f"{sub['nt']} of {nt} instances (every {nt // sub['nt']}th)"
        cpu = cpu_record(steps / secs, "instance-steps/s",
                         f"{what}, {runs} run(s): {int(steps)} accepted steps in {secs:.2f} s, FP64 OpenMP oracle")
        mine = out if cpu_idx is None else {k: (v[..., cpu_idx] if isinstance(v, np.ndarray) and v.ndim and v.shape[-1] == nt else v)
                                            for k, v in out.items()}
        if prec == 0:
            parity = {"bitwise_equal_to_oracle": bool(bitwise(mine, ref, ("X", "T", "nacc", "nrej", "status"))),
                      "instances_compared": int(sub["nt"])}
        else:
            scale = np.maximum(np.abs(ref["X"]), 1e-3)
            parity = {"max_rel_state_error_vs_fp64_oracle": float(np.max(np.abs(mine["X"] - ref["X"]) / scale)),
                      "median_rel_state_error": float(np.median(np.abs(mine["X"] - ref["X"]) / scale)),
                      "accepted_step_count_mismatches": int(np.sum(mine["nacc"] != ref["nacc"])),
                      "instances_compared": int(sub["nt"])}
    line = contract_line(args, value, "instance-steps/s", res["wall_ms_per_step"],
                         {"workload": name, "instances": nt, **desc, "kernel_form": best, "launch": res["config"],
                          "accepted_steps_per_solve": accepted, "rejected_steps_per_solve": attempted - accepted,
                          "time_sliced_steps": sl or None},
                         {"bound": "fp64_valu" if prec == 0 else "fp32_valu", "engine": ENGINE_NOTE if prec == 0 else ENGINE_NOTE_F32,
                          "achieved": achieved,
                          "peak": FP64_PEAK_TFLOPS if prec == 0 else FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": achieved / (FP64_PEAK_TFLOPS if prec == 0 else FP32_PEAK_TFLOPS),
                          "traffic": traffic_for(name) if prec == 0 else None, "flops_per_attempted_step": fpa, "kernel_ms": res["kernel_ms"]},
                         cpu, parity)
    line["forms"] = {k: {"config": v["config"], "wall_ms_per_step": v["wall_ms_per_step"], "kernel_ms": v["kernel_ms"],
                         "instance_steps_per_s": accepted / (v["wall_ms_per_step"] * 1e-3)} for k, v in forms.items()}
    if "plain" in forms and "persistent" in forms:
        line["persistent_over_plain"] = forms["plain"]["kernel_ms"] / forms["persistent"]["kernel_ms"]
    if "plain" in forms and "relaunch" in forms:
        line["relaunch_over_plain"] = forms["plain"]["kernel_ms"] / forms["relaunch"]["kernel_ms"]
    if "plain" in forms and "sorted" in forms:
        line["sorted_over_plain"] = forms["plain"]["kernel_ms"] / forms["sorted"]["kernel_ms"]
    return line


# --------------------------------------------------------------------------- adjoint
def adjoint(args, mpgos, W, pyoracle):
    from mpgos.enums import Algorithms, ListOfVariables as V
    from mpgos.tuning import DeviceSnapshot
    alg = Algorithms.RK4 if args.algorithm == "rk4" else Algorithms.RKCK45
    spec = W.config3_duffing_adjoint(nt=args.nt or 65536, algorithm=alg)
    nt = spec["nt"]
    K = args.checkpoint_interval
    ktable = None
    s = mpgos.ProblemSolver(spec["system"], nt, alg)
    try:
        W.load(s, spec)
        snap = DeviceSnapshot(s)
        lam = np.zeros((3, nt))
        lam[2] = 1.0                                    # dL/dq(H): d/dq of the quadrature state
        s.Solve()
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        out = W.read(s)
        steps_needed = int(out["nacc"].max())
        snap.restore()
        snap.sync()
        if K == 0:                                      # checkpoint interval from the autotuner
            from mpgos.tuning import autotune_adjoint
            K, ktable = autotune_adjoint(s, lam, steps_needed, candidates=(4, 6, 8, 12, 16, 24, 32, 64, 96, 128), log=log)
        cap = steps_needed // K + 2
        fwd, rec, rev = [], [], []
        for i in range(args.warmup + args.steps):
            snap.restore()
            s.Solve()
            snap.sync()
            f_ms = s.last_solve_ms()
            snap.restore()
            s.adjoint_record(K, cap)
            snap.sync()
            r_ms = s.last_solve_ms()
            s.host_array(V.AdjointState)[...] = lam
            s.SynchroniseFromHostToDevice(V.AdjointState)
            s.adjoint_reverse()
            snap.sync()
            b_ms = s.last_solve_ms()
            if i >= args.warmup:
                fwd.append(f_ms); rec.append(r_ms); rev.append(b_ms)
        snap.close()
    finally:
        s.close()
    f, r, b = float(np.mean(fwd)), float(np.mean(rec)), float(np.mean(rev))
    accepted = int(out["nacc"].sum())
    attempted = accepted + int(out["nrej"].sum())
    fpa = flops_per_attempt(pyoracle, W, spec, np.arange(0, nt, nt // 16))
    achieved = fpa * attempted / (f * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        ref, steps, secs, runs = timed_oracle(lambda: pyoracle.run(spec["system"], spec, nthreads=threads_here()))
        cpu = cpu_record(steps / secs, "instance-steps/s",
                         f"forward pass of all {nt} instances, {runs} run(s): {int(steps)} accepted steps in {secs:.2f} s")
    line = contract_line(args, accepted / (f * 1e-3), "instance-steps/s", f,
                         {"workload": f"duffing_adjoint_{args.algorithm}_{nt}", "instances": nt,
                          "algorithm": args.algorithm.upper(), "checkpoint_interval": K,
                          "accepted_steps_per_solve": accepted},
                         {"bound": "fp64_valu", "engine": ENGINE_NOTE, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                          "flops_per_attempted_step": fpa, "kernel_ms": f},
                         cpu, None)
    line["adjoint"] = {"forward_ms": f, "record_ms": r, "reverse_ms": b, "adjoint_over_forward": (r + b) / f,
                       "reverse_over_forward": b / f, "gradient_run_extra_over_forward": (r + b - f) / f,
                       "checkpoint_interval": K,
                       "checkpoint_interval_from": "autotuner" if ktable else "command line"}
    if ktable:
        line["adjoint"]["checkpoint_autotune"] = ktable
    return line


# --------------------------------------------------------------------------- per-block ring
RING_FLOPS_NOTE = ("per system-step: 6 stages x 2 x U^2 (coupling) + U x 404 (unit right-hand sides with cos at 38, "
                   "stage combinations, error, tolerance) = 889,856 at U = 256")


def ring(args, mpgos, W, pyoracle):
    from mpgos.perblock import PBOption, PBProblemSolver
    ns = args.nt or 4096
    U = 256
    spec = W.config4_ring(ns=ns, ups=U, periods=4)
    # (coupling kernel, units per lane of the dense-contraction kernel)
    all_forms = {"mfma": (2, 4), "mfma2": (2, 2), "lane": (1, 4)}
    kernels = all_forms if args.form == "auto" else {args.form: all_forms[args.form]}
    forms = {}
    for name, (k, upl) in kernels.items():
        s = PBProblemSolver("ring_duffing", ns, U, spec["algorithm"])
        try:
            s.SolverOption(PBOption.CouplingKernel, k)
            s.SolverOption(PBOption.MfmaUnitsPerLane, upl)
            times = []
            for i in range(args.warmup + args.steps):
                W.load_pb(s, spec)
                s.Solve()
                s.SynchroniseDevice()
                if i >= args.warmup:
                    times.append(s.last_solve_ms())
            s.SynchroniseFromDeviceToHost(0)
            s.InsertSynchronisationPoint()
            s.SynchroniseSolver()
            forms[name] = {"kernel_ms": float(np.mean(times)), "out": W.read_pb(s)}
        finally:
            s.close()
    best = min(forms, key=lambda k: forms[k]["kernel_ms"])
    res = forms[best]
    out = res["out"]
    accepted = int(out["nacc"].sum())
    attempted = accepted + int(out["nrej"].sum())
    flops_step = 6 * 2 * U * U + U * 404
    achieved = flops_step * attempted / (res["kernel_ms"] * 1e-3) / 1e12
    cpu = parity = None
    if not args.no_cpu:
        idx = np.arange(0, ns, max(1, ns // 64))
        sub = W.subset_pb(spec, idx)
        ref = pyoracle.run_pb("ring_duffing", sub, nthreads=threads_here())
        cpu = cpu_record(float(ref["nacc"].sum() / ref["seconds"]), "system-steps/s",
                         f"{len(idx)} of {ns} systems (every {ns // len(idx)}th): "
                         f"{int(ref['nacc'].sum())} accepted system-steps in {ref['seconds']:.2f} s")
        mine = W.subset_pb_out(out, idx, U)
        parity = {"bitwise_equal_to_oracle": bool(bitwise(mine, ref, ("X", "T", "nacc", "nrej", "status"))),
                  "systems_compared": len(idx)}
        if len(forms) > 1:
            parity["mfma_equals_lane_kernel_bitwise"] = bool(bitwise(forms["mfma"]["out"], forms["lane"]["out"],
                                                                     ("X", "T", "nacc", "nrej")))
    line = contract_line(args, accepted / (res["kernel_ms"] * 1e-3), "system-steps/s", res["kernel_ms"],
                         {"workload": f"ring_duffing_{ns}x{U}", "systems": ns, "units_per_system": U,
                          "algorithm": "RKCK45", "rtol": 1e-9, "atol": 1e-9, "periods": 4, "kernel_form": best,
                          "accepted_system_steps": accepted},
                         {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic_for(f"ring_duffing_{ns}x{U}"),
                          "flops_per_attempted_step": flops_step, "flops_note": RING_FLOPS_NOTE,
                          "kernel_ms": res["kernel_ms"]},
                         cpu, parity)
    line["forms"] = {k: {"kernel_ms": v["kernel_ms"], "system_steps_per_s": accepted / (v["kernel_ms"] * 1e-3)}
                     for k, v in forms.items()}
    if not args.no_adjoint:
        line["adjoint"] = ring_adjoint(args, W, ns, U, kernels[best][0])
    return line


def ring_adjoint(args, W, ns, U, kernel):
    """Cost of the per-block discrete adjoint of the ring: the same rings with a
    quadrature unit state q' = x^2 (the loss is a combination of q at the
    ends of the last forcing period), gradient w.r.t. the per-system forcing
    amplitude B.  Forward Solve, record pass and reverse pass of the same
    system, each timed with HIP events on the solver's stream."""
    from mpgos.perblock import PB, PBOption, PBProblemSolver
    spec = W.config4_ring(ns=ns, ups=U, periods=4, system="ring_duffing_q")
    K = args.checkpoint_interval
    ktable = None
    s = PBProblemSolver("ring_duffing_q", ns, U, spec["algorithm"])
    fwd, rec, rev = [], [], []
    try:
        s.SolverOption(PBOption.CouplingKernel, kernel)
        W.load_pb(s, spec)
        s.solve_blocking()
        nmax = int(W.read_pb(s)["nacc"].max())
        lam = np.zeros((3, ns * U))
        lam[2] = 1.0 / (2.0 * np.pi * U)                   # d(mean over units of q(end))/dq / period
        if K == 0:                                          # checkpoint interval from the autotuner
            from mpgos.tuning import autotune_checkpoint_interval

            def record(k):
                W.load_pb(s, spec)
                s.adjoint_record(k, nmax // k + 2, nmax + 2)
                s.SynchroniseDevice()
                return s.last_solve_ms()

            def reverse():
                s.host_array(PB.AdjointState)[...] = lam
                s.SynchroniseFromHostToDevice(PB.AdjointState)
                s.adjoint_reverse()
                s.SynchroniseDevice()
                return s.last_solve_ms()
            K, ktable = autotune_checkpoint_interval(record, reverse, candidates=(8, 16, 32, 64), log=log)
        for i in range(args.warmup + args.steps):
            W.load_pb(s, spec)
            s.Solve()
            s.SynchroniseDevice()
            f_ms = s.last_solve_ms()
            W.load_pb(s, spec)
            s.adjoint_record(K, nmax // K + 2, nmax + 2)
            s.SynchroniseDevice()
            r_ms = s.last_solve_ms()
            s.host_array(PB.AdjointState)[...] = lam
            s.SynchroniseFromHostToDevice(PB.AdjointState)
            s.adjoint_reverse()
            s.SynchroniseDevice()
            b_ms = s.last_solve_ms()
            log(f"ring adjoint rep {i}: forward {f_ms:.1f} ms, record {r_ms:.1f} ms, reverse {b_ms:.1f} ms")
            if i >= args.warmup:
                fwd.append(f_ms); rec.append(r_ms); rev.append(b_ms)
        s.SynchroniseFromDeviceToHost(PB.SystemParameterGradient)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        gB = s.host_array(PB.SystemParameterGradient)[0].copy()
    finally:
        s.close()
    f, r, b = float(np.mean(fwd)), float(np.mean(rec)), float(np.mean(rev))
    return {"system": "ring_duffing_q (unit state x, v, q with q' = x^2)", "gradient": "dL/dB per system",
            "checkpoint_interval": K, "checkpoint_interval_from": "autotuner" if ktable else "command line",
            "forward_ms": f, "record_ms": r, "reverse_ms": b,
            "adjoint_over_forward": (r + b) / f, "reverse_over_forward": b / f,
            "gradient_run_extra_over_forward": (r + b - f) / f,
            "gradient_finite": bool(np.isfinite(gB).all()), **({"checkpoint_autotune": ktable} if ktable else {})}


# --------------------------------------------------------------------------- the rest of the metric
def headline_extras(args, mpgos, W, pyoracle):
    """The other half of the metric, carried by the default line: the adjoint
    over forward cost of the Duffing ensemble (RK4 and RKCK45, autotuned
    checkpoint interval) and the per-block ring of 256 coupled units (MFMA
    coupling kernel, its CPU sample and parity, and its adjoint), each with
    fewer timed repetitions than its own --workload run."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.nt, sub.precision, sub.checkpoint_interval = 3, 1, 0, "f64", 0
    adj = {}
    for alg in ("rk4", "rkck45"):
        sub.algorithm = alg
        log(f"extras: adjoint {alg}")
        l = adjoint(sub, mpgos, W, pyoracle)
        a = dict(l["adjoint"])
        a.pop("checkpoint_autotune", None)
        adj[alg] = {"workload": l["config"]["workload"], "forward_instance_steps_per_s": l["value"],
                    "forward_frac_fp64_peak": l["roofline"]["frac"], **a,
                    "cpu_baseline": l["cpu_baseline"], "speedup_vs_cpu": l.get("speedup_vs_cpu")}
    sub.form = "mfma"
    sub.no_adjoint = False
    log("extras: ring")
    r = ring(sub, mpgos, W, pyoracle)
    ring_obj = {"workload": r["config"]["workload"], "value": r["value"], "unit": r["unit"],
                "kernel_ms": r["roofline"]["kernel_ms"], "roofline": {k: r["roofline"][k] for k in ("bound", "achieved", "peak", "unit", "frac")},
                "cpu_baseline": r["cpu_baseline"], "speedup_vs_cpu": r.get("speedup_vs_cpu"), "parity": r["parity"],
                "adjoint": {k: v for k, v in r.get("adjoint", {}).items() if k != "checkpoint_autotune"}}
    return adj, ring_obj


# --------------------------------------------------------------------------- common
def traffic_for(name):
    tfile = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def contract_line(args, value, unit, ms_per_step, config, roofline, cpu, parity):
    line = {
        "metric": "ODE instance-steps/s at rtol 1e-9 vs FP64 oracle; adjoint/forward cost",
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32", "f32k": "f32 (Kahan-compensated state and time)"}[args.precision],
        "data": "synthetic (seeded SplitMix64, seed 20261202)",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if cpu:
        line["speedup_vs_cpu"] = value / cpu["value"]
    return line


def copy_calibration(mib):
    from mpgos.tuning import hip
    h = hip()
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if h.hipMalloc(ctypes.byref(a), n) != 0 or h.hipMalloc(ctypes.byref(b), n) != 0:
        raise MemoryError("hipMalloc failed")
    h.hipDeviceSynchronize.argtypes = []
    for _ in range(3):
        if h.hipMemcpy(b, a, n, 3) != 0:
            raise RuntimeError("copy failed")
    h.hipDeviceSynchronize()
    h.hipFree(a)
    h.hipFree(b)
    print(json.dumps({"copy_calibration": {"bytes_per_copy": n, "copies": 3,
                                           "note": "device-to-device hipMemcpy: reads and writes bytes_per_copy each"}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lorenz",
                    choices=["lorenz", "duffing_rk4", "duffing_uneven", "duffing_adjoint", "ring"])
    ap.add_argument("--nt", type=int, default=0, help="instances (systems for ring); 0 = the workload's size")
    ap.add_argument("--form", default="auto",
                    help="auto | plain | persistent | relaunch | sorted (per-thread; duffing_uneven also: single = one "
                         "plain launch, relaunch = device-chained launches of --slice-steps); auto | mfma | mfma2 | lane (ring; mfma2 = two units per lane)")
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "f32k"])
    ap.add_argument("--algorithm", default="rk4", choices=["rk4", "rkck45"], help="duffing_adjoint only")
    ap.add_argument("--checkpoint-interval", type=int, default=0,
                    help="adjoint checkpoint interval K; 0 = picked by the autotuner")
    ap.add_argument("--slice-steps", type=int, default=2000000,
                    help="duffing_uneven: attempted steps per instance per launch of the time-sliced solve")
    ap.add_argument("--solo", type=int, default=0,
                    help="duffing_uneven, persistent form: longest instances run alone in a workgroup")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--no-adjoint", action="store_true", help="ring: skip the per-block adjoint timing")
    ap.add_argument("--headline-only", action="store_true",
                    help="lorenz: print the Lorenz line without the adjoint and ring figures of the default run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solver path has no inter-GPU work "
              "(instances and systems shard by index range with zero exchange); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import mpgos
    from mpgos import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if args.workload in ("lorenz", "duffing_rk4", "duffing_uneven"):
        line = per_thread(args, mpgos, W, pyoracle)
        if args.workload == "lorenz" and not args.headline_only and args.precision == "f64":
            line["adjoint"], line["ring"] = headline_extras(args, mpgos, W, pyoracle)
    elif args.workload == "duffing_adjoint":
        line = adjoint(args, mpgos, W, pyoracle)
    else:
        line = ring(args, mpgos, W, pyoracle)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
