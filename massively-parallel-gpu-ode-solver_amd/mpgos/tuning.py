"""Timing helpers and the launch-shape autotuner.

DeviceSnapshot keeps device-side copies of a solver's input buffers so repeated
Solve() calls start from identical inputs without host traffic (inputs stay
resident in HBM).  autotune() times candidate launch configurations on the
caller's real workload and returns the fastest:
  kernel form      plain (one lane per instance), persistent (work queue
                   with lane compaction), relaunch (one lane per instance,
                   relaunched every RelaunchSteps steps with the unfinished
                   instances repacked) or sorted (a pilot launch, then the
                   instances ordered by estimated remaining work),
  ThreadsPerBlock  64 .. 512,
  LaunchVariant    register budget (waves-per-EU hint 1, 2, 4, 8),
  PersistentRoundSteps  attempted steps between compactions.
Results are bit-identical across candidates (the arithmetic of an instance
does not depend on its lane), so the choice is purely about speed.  The
winning configuration can be cached in a JSON file keyed by system,
algorithm, precision and instance count.
autotune_checkpoint_interval() does the same for the checkpoint interval K of
the discrete adjoint (record + reverse time; the gradients do not depend on
K), for per-thread and per-block solvers alike.
"""
import ctypes
import itertools
import json
import os
import time

import numpy as np

from ._lib import hip_runtime_path
from .enums import ListOfSolverOptions as O, ListOfVariables as V

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        h = ctypes.CDLL(hip_runtime_path())
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _HIP = h
    return _HIP


class DeviceSnapshot:
    """Device copies of the input variables of `solver` (taken now); restore()
    copies them back on the solver's stream (device to device)."""

    INPUTS = (V.TimeDomain, V.ActualState, V.ActualTime, V.ControlParameters, V.Accessories,
              V.IntegerAccessories, V.DenseIndex)

    def __init__(self, solver, variables=INPUTS):
        self.solver = solver
        self.saved = []
        h = hip()
        for var in variables:
            ptr, n = solver.device_ptr(var)
            if not ptr or n == 0:
                continue
            dst = ctypes.c_void_p()
            if h.hipMalloc(ctypes.byref(dst), n) != 0:
                raise MemoryError("hipMalloc failed for the input snapshot")
            if h.hipMemcpyAsync(dst, ptr, n, 3, solver.stream()) != 0:
                raise RuntimeError("snapshot copy failed")
            self.saved.append((ptr, dst, n))
        self.sync()

    def sync(self):
        if hip().hipStreamSynchronize(self.solver.stream()) != 0:
            raise RuntimeError("hipStreamSynchronize failed")

    def restore(self):
        for ptr, src, n in self.saved:
            if hip().hipMemcpyAsync(ptr, src, n, 3, self.solver.stream()) != 0:
                raise RuntimeError("restore copy failed")

    def close(self):
        self.sync()
        for _, src, _ in self.saved:
            hip().hipFree(src)
        self.saved = []


def apply(solver, cfg):
    solver.SolverOption(O.KernelMode, cfg.get("mode", 0))
    if "block" in cfg:
        solver.SolverOption(O.ThreadsPerBlock, cfg["block"])
    solver.SolverOption(O.LaunchVariant, cfg.get("variant", 0))
    solver.SolverOption(O.PersistentRoundSteps, cfg.get("round", 32))
    solver.SolverOption(O.PersistentWavesPerSimd, cfg.get("waves_per_simd", 0))
    solver.SolverOption(O.QueueOrder, cfg.get("order", 0))
    solver.SolverOption(O.PersistentSoloInstances, cfg.get("solo", 0))
    solver.SolverOption(O.RelaunchSteps, cfg.get("relaunch", 300))
    solver.SolverOption(O.RelaunchLaunches, cfg.get("launches", 16))
    if "graph" in cfg:                          # else the solver's own GraphCapture setting stays
        solver.SolverOption(O.GraphCapture, cfg["graph"])


def time_config(solver, snap, cfg, repeats=2, warmup=1):
    """Mean wall time (ms) of restore + Solve for one configuration; also the
    mean kernel time from the solver's HIP events."""
    apply(solver, cfg)
    for _ in range(warmup):
        snap.restore()
        solver.Solve()
    snap.sync()
    kern = []
    t0 = time.perf_counter()
    for _ in range(repeats):
        snap.restore()
        solver.Solve()
        snap.sync()
        kern.append(solver.last_solve_ms())
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / repeats, float(np.mean(kern))


def default_candidates():
    plain = [{"mode": 0, "block": b, "variant": v} for b, v in itertools.product((64, 128, 256), (0, 2))]
    pers = [{"mode": 1, "block": b, "variant": v, "round": r}
            for b, v, r in itertools.product((256, 512), (0, 2), (8, 32, 128))]
    # (attempted steps per launch, budgeted launches): launches beyond the
    # longest instance's need are empty dispatches, too few leave the tail to
    # the last (unbudgeted) launch without repacking
    shapes = [(r, n) for r in (100, 200) for n in (8, 12, 16, 32)] + [(50, 32), (400, 32)]
    relaunch = [{"mode": 2, "block": b, "variant": v, "relaunch": r, "launches": n}
                for b, v, (r, n) in itertools.product((64, 256), (0, 2), shapes)]
    # sorted form: pilot launch of r attempted steps, then the instances
    # ordered by estimated remaining work
    srt = [{"mode": 3, "block": b, "variant": v, "relaunch": r}
           for b, v, r in itertools.product((64, 256), (0, 2), (200, 400, 600, 800))]
    return plain + pers + relaunch + srt


def autotune(solver, candidates=None, repeats=2, cache=None, key=None, log=None):
    """Time every candidate on the solver's current inputs (restored between
    runs) and leave the fastest applied.  Returns (best_cfg, table)."""
    if cache and key and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if key in known:
            apply(solver, known[key]["config"])
            return known[key]["config"], known[key]["table"]
    snap = DeviceSnapshot(solver)
    table = []
    try:
        for cfg in candidates or default_candidates():
            try:
                wall, kern = time_config(solver, snap, cfg, repeats)
            except Exception as e:  # a shape the device rejects (e.g. LDS limit)
                if log:
                    log(f"autotune: skip {cfg}: {e}")
                continue
            table.append({"config": cfg, "wall_ms": wall, "kernel_ms": kern})
            if log:
                log(f"autotune: {cfg} wall {wall:.3f} ms kernel {kern:.3f} ms")
        snap.restore()
        snap.sync()
    finally:
        snap.close()
    best = min(table, key=lambda r: r["kernel_ms"])["config"]
    apply(solver, best)
    if cache and key:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[key] = {"config": best, "table": table}
        with open(cache, "w") as f:
            json.dump(known, f, indent=1)
    return best, table


def autotune_checkpoint_interval(run_record, run_reverse, candidates=(8, 16, 32, 64, 128, 256), repeats=1, log=None):
    """Checkpoint interval K of the discrete adjoint with the smallest
    record + reverse time.  `run_record(K)` performs a record pass from the
    solver's inputs and returns its kernel time in ms, `run_reverse()` the
    reverse pass (with dL/dx(end) already in place) and returns its time.  The
    gradients are identical bit for bit for every K, so the choice only trades
    recomputation and memory for speed.  Returns (K, table)."""
    table = []
    for K in candidates:
        rec, rev = [], []
        try:
            for _ in range(max(1, repeats)):
                rec.append(run_record(K))
                rev.append(run_reverse())
        except Exception as e:  # e.g. an allocation the device cannot hold
            if log:
                log(f"autotune K: skip {K}: {e}")
            continue
        row = {"K": K, "record_ms": float(np.mean(rec)), "reverse_ms": float(np.mean(rev))}
        row["total_ms"] = row["record_ms"] + row["reverse_ms"]
        table.append(row)
        if log:
            log(f"autotune K={K}: record {row['record_ms']:.2f} ms reverse {row['reverse_ms']:.2f} ms")
    if not table:
        raise RuntimeError("autotune_checkpoint_interval: no candidate ran")
    return min(table, key=lambda r: r["total_ms"])["K"], table


def autotune_adjoint(solver, lam_T, max_steps, candidates=(4, 8, 16, 32, 64, 128, 256), repeats=1, log=None):
    """autotune_checkpoint_interval for a per-thread ProblemSolver: every record
    pass starts from the inputs held when called (device snapshot); `lam_T` is
    dL/dx(T) (SD x NT) and `max_steps` the largest accepted-step count of an
    instance (the checkpoint capacity is sized from it).  Leaves the solver's
    inputs restored.  Returns (K, table)."""
    snap = DeviceSnapshot(solver)

    def record(K):
        snap.restore()
        solver.adjoint_record(K, max_steps // K + 2)
        snap.sync()
        return solver.last_solve_ms()

    def reverse():
        solver.host_array(V.AdjointState)[...] = lam_T
        solver.SynchroniseFromHostToDevice(V.AdjointState)
        solver.adjoint_reverse()
        snap.sync()
        return solver.last_solve_ms()
    try:
        return autotune_checkpoint_interval(record, reverse, candidates, repeats, log)
    finally:
        snap.restore()
        snap.sync()
        snap.close()
