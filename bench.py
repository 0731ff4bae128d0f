"""Benchmark of the MI355X MPGOS solver path.

Workload (one GPU): the Lorenz ensemble, 262144 instances, rho ~ U[0, 21],
x0 = (10, 10, 10) + N(0, 0.01), adaptive RKCK45 at rtol = atol = 1e-10 over
t in [0, 10], FP64.  One "step" is one Solve() over the whole batch with the
inputs already resident in HBM: a device-to-device reset of the initial state
followed by the integration kernel.  The metric is accepted instance-steps per
second (one accepted Runge-Kutta step of one instance).

Printed (rank 0, one JSON line): the contract fields plus
  roofline      -- algorithmic FP64 FLOP/s of the integration kernel (FLOPs per
                   attempted step from the oracle's op count x attempted steps,
                   divided by the kernel time measured with HIP events on the
                   solver's stream) against the 78.6 TFLOP/s FP64 peak;
  cpu_baseline  -- the FP64 OpenMP oracle on the host cores, same workload;
  parity        -- the GPU result compared bit for bit with that oracle run;
  forms         -- the plain kernel in the reference's launch shape and the
                   configuration chosen by the autotuner (mpgos/tuning.py).
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
HBM_PEAK_GBS = 8000.0


def time_form(mpgos, spec, cfg, steps, warmup):
    """K timed steps (device-side input reset + Solve) of one launch configuration."""
    from mpgos import workloads as W
    from mpgos.tuning import DeviceSnapshot, apply
    s = mpgos.ProblemSolver(spec["system"], spec["nt"], spec["algorithm"])
    try:
        W.load(s, spec)
        apply(s, cfg)
        snap = DeviceSnapshot(s)
        for _ in range(warmup):
            snap.restore()
            s.Solve()
        snap.sync()
        kernel_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            snap.restore()
            s.Solve()
            snap.sync()
            kernel_ms.append(s.last_solve_ms())
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


def tune(mpgos, spec):
    from mpgos import workloads as W
    from mpgos.tuning import autotune
    s = mpgos.ProblemSolver(spec["system"], spec["nt"], spec["algorithm"])
    try:
        W.load(s, spec)
        best, table = autotune(s, log=lambda m: print(m, file=sys.stderr))
    finally:
        s.close()
    return best, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nt", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solver path has no inter-GPU work "
              "(instances shard by index range with zero exchange); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import mpgos
    from mpgos import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

    spec = W.config1_lorenz(nt=args.nt)
    forms = {"plain": time_form(mpgos, spec, {"mode": 0, "block": 64}, args.steps, args.warmup)}
    if not args.no_tune:
        best_cfg, table = tune(mpgos, spec)
        forms["autotuned"] = time_form(mpgos, spec, best_cfg, args.steps, args.warmup)
    else:
        forms["persistent"] = time_form(mpgos, spec, {"mode": 1, "block": 256, "round": 32}, args.steps, args.warmup)
    best = min(forms, key=lambda k: forms[k]["wall_ms_per_step"])
    res = forms[best]
    out = res["out"]
    accepted = int(out["nacc"].sum())
    attempted = accepted + int(out["nrej"].sum())
    value = accepted / (res["wall_ms_per_step"] * 1e-3)

    # algorithmic FLOPs per attempted step: oracle op count over a sample
    import pyoracle
    sample = W.subset(spec, np.arange(0, args.nt, max(1, args.nt // 256)))
    fl = pyoracle.run("lorenz", sample, count_flops=True)
    flops_per_attempt = fl["flops"] / float(fl["nacc"].sum() + fl["nrej"].sum())
    flops_per_launch = flops_per_attempt * attempted
    achieved_tflops = flops_per_launch / (res["kernel_ms"] * 1e-3) / 1e12

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "lorenz_rkck45_262144_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    parity = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        ref = pyoracle.run("lorenz", spec, nthreads=threads)
        cpu = {"value": float(ref["nacc"].sum() / ref["seconds"]), "unit": "instance-steps/s",
               "cores": threads, "kind": "port",
               "sample": f"full workload ({args.nt} instances, {int(ref['nacc'].sum())} accepted steps), "
                         f"{ref['seconds']:.2f} s, FP64 OpenMP oracle schedule(dynamic,64)"}
        same = all(np.array_equal(np.asarray(out[k]).view(np.uint8), np.asarray(ref[k]).view(np.uint8))
                   for k in ("X", "T", "nacc", "nrej", "status"))
        parity = {"bitwise_equal_to_oracle": bool(same), "instances": args.nt}

    line = {
        "metric": "ODE instance-steps/s at rtol 1e-9 vs FP64 oracle; adjoint/forward cost",
        "value": value,
        "unit": "instance-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["wall_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64, seed 20261202, stream 1)",
        "config": {"workload": f"lorenz_rkck45_{args.nt}", "system": "lorenz", "algorithm": "RKCK45",
                   "instances": args.nt, "rtol": 1e-10, "atol": 1e-10, "t_end": 10.0,
                   "kernel_form": best, "launch": res["config"], "accepted_steps_per_solve": accepted,
                   "rejected_steps_per_solve": attempted - accepted},
        "roofline": {"bound": "fp64-valu", "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "flops_per_attempted_step": flops_per_attempt, "kernel_ms": res["kernel_ms"]},
        "cpu_baseline": cpu,
        "parity": parity,
        "forms": {k: {"config": v["config"], "wall_ms_per_step": v["wall_ms_per_step"], "kernel_ms": v["kernel_ms"],
                      "instance_steps_per_s": accepted / (v["wall_ms_per_step"] * 1e-3)} for k, v in forms.items()},
    }
    if cpu:
        line["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
