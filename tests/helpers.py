"""Shared helpers of the parity tests: run one spec on the GPU and on the oracle."""
import numpy as np

from mpgos import ListOfSolverOptions as O, ListOfVariables as V, ProblemSolver, Precision
from mpgos import workloads as W


def run_gpu(spec, precision=Precision.DOUBLE, kernel_mode=0, block=None, variant=0, graph=0, solves=1,
            waves_per_simd=0, slice_steps=0, queue_order=0, round_steps=0, solo=0, relaunch_steps=0,
            relaunch_launches=None):
    s = ProblemSolver(spec["system"], spec["nt"], spec["algorithm"], precision, ndo=spec.get("ndo", 0))
    try:
        W.load(s, spec)
        s.SolverOption(O.KernelMode, kernel_mode)
        s.SolverOption(O.LaunchVariant, variant)
        s.SolverOption(O.GraphCapture, graph)
        s.SolverOption(O.PersistentWavesPerSimd, waves_per_simd)
        s.SolverOption(O.QueueOrder, queue_order)
        s.SolverOption(O.PersistentSoloInstances, solo)
        if relaunch_steps:
            s.SolverOption(O.RelaunchSteps, relaunch_steps)
        if relaunch_launches is not None:
            s.SolverOption(O.RelaunchLaunches, relaunch_launches)
        if round_steps:
            s.SolverOption(O.PersistentRoundSteps, round_steps)
        if block:
            s.SolverOption(O.ThreadsPerBlock, block)
        outs = []
        for _ in range(solves):
            if slice_steps:
                slices = 1
                while s.solve_slice(slice_steps) > 0:
                    slices += 1
                s.SynchroniseFromDeviceToHost(V.All)
                s.InsertSynchronisationPoint()
                s.SynchroniseSolver()
                out = W.read(s)
                out["_slices"] = slices
                outs.append(out)
            else:
                s.solve_blocking()
                outs.append(W.read(s))
        return outs if solves > 1 else outs[0]
    finally:
        s.close()


def mask_unwritten(out):
    """Zero the event-log and dense-output slots past each instance's count
    (their content is unspecified)."""
    out = dict(out)
    if "evlog" in out and out["evlog"].size:
        cap = out["evlog"].shape[0]
        valid = np.arange(cap)[:, None] < np.minimum(out["nev"], cap)[None, :]
        out["evlog"] = np.where(valid[:, None, :], out["evlog"], 0.0)
        out["evlogidx"] = np.where(valid, out["evlogidx"], 0)
    if "DOt" in out and out["DOt"].size:
        ndo = out["DOt"].shape[0]
        valid = np.arange(ndo)[:, None] < np.minimum(out["DOidx"], ndo)[None, :]
        out["DOt"] = np.where(valid, out["DOt"], 0.0)
        out["DOx"] = np.where(valid[:, None, :], out["DOx"], 0.0)
    return out


def assert_bitwise(gpu, ref, keys=("X", "T", "ACC", "ACCi", "nacc", "nrej", "status", "nev", "DOidx"), label=""):
    gpu, ref = mask_unwritten(gpu), mask_unwritten(ref)
    for k in keys:
        if k not in gpu or k not in ref:
            continue
        a, b = np.asarray(gpu[k]), np.asarray(ref[k])
        assert a.shape == b.shape, f"{label} {k}: shape {a.shape} vs {b.shape}"
        if a.dtype.kind == "f":
            same = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))
        else:
            same = a == b
        if not same.all():
            bad = np.argwhere(~same)
            i = tuple(bad[0])
            raise AssertionError(f"{label} {k}: {len(bad)} of {a.size} differ; first at {i}: gpu={a[i]!r} oracle={b[i]!r}")
