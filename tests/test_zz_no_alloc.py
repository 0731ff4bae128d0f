"""The allocation-free Solve (mpgos_pt_solve_noalloc) against the Solve that
allocates its relaunch buffers, through the C ABI.

Two calls in a row on the static scratch memory, the second with a smaller
input (fewer instances, and a different system and algorithm), each compared
bit for bit with mpgos_pt_solve in the relaunch form and with the oracle; a
handle too large for the scratch memory is refused before anything runs.
"""
import numpy as np
import pytest

from helpers import assert_bitwise, run_gpu
from mpgos import ListOfSolverOptions as O, ListOfVariables as V, ProblemSolver
from mpgos import workloads as W
from mpgos._lib import MPGOSError

pytestmark = pytest.mark.gpu

KEYS = ("X", "T", "ACC", "ACCi", "nacc", "nrej", "status", "nev", "evlog", "evlogidx")


def run_noalloc(spec, relaunch_steps, relaunch_launches, block=64):
    s = ProblemSolver(spec["system"], spec["nt"], spec["algorithm"], ndo=spec.get("ndo", 0))
    try:
        W.load(s, spec)
        s.SolverOption(O.RelaunchSteps, relaunch_steps)
        s.SolverOption(O.RelaunchLaunches, relaunch_launches)
        s.SolverOption(O.ThreadsPerBlock, block)
        s.solve_noalloc()
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        return W.read(s)
    finally:
        s.close()


def test_noalloc_matches_allocating_solve_twice(gpu, oracle):
    # first call: 4096 Lorenz instances (RKCK45), many suspensions per instance
    big = W.config1_lorenz(nt=4096)
    print("noalloc call 1: lorenz RKCK45, 4096 instances", flush=True)
    out1 = run_noalloc(big, relaunch_steps=64, relaunch_launches=12)
    ref1 = run_gpu(big, kernel_mode=2, block=64, relaunch_steps=64, relaunch_launches=12)
    assert_bitwise(out1, ref1, keys=KEYS, label="noalloc call 1 vs relaunch Solve")
    assert_bitwise(out1, oracle.run("lorenz", big), keys=KEYS, label="noalloc call 1 vs oracle")
    assert int(out1["nacc"].max()) > 64 * 12, "the last (unbudgeted) launch must have run"
    # second call, smaller input: 600 uneven Duffing instances with events (the
    # scratch memory still holds the first call's records)
    full = W.config2_duffing_uneven(nt=131072, evcap=8)
    small = W.subset(full, np.arange(0, 131072, 131072 // 600)[:600])
    small["TD"] = small["TD"].copy()
    small["TD"][1] = np.minimum(small["TD"][1], 300.0)
    print(f"noalloc call 2: duffing_eps RKCK45 + events, {small['nt']} instances", flush=True)
    out2 = run_noalloc(small, relaunch_steps=100, relaunch_launches=5, block=128)
    ref2 = run_gpu(small, kernel_mode=2, block=128, relaunch_steps=100, relaunch_launches=5)
    assert_bitwise(out2, ref2, keys=KEYS, label="noalloc call 2 vs relaunch Solve")
    assert_bitwise(out2, oracle.run("duffing_eps", small), keys=KEYS, label="noalloc call 2 vs oracle")
    assert int(out2["nev"].max()) > 8


def test_noalloc_rk4_and_partial_activity(gpu, oracle):
    spec = W.config0_duffing_rk4(nt=1000)
    spec["TD"] = spec["TD"].copy()
    spec["TD"][1] = np.full(spec["nt"], 4 * np.pi)
    s = ProblemSolver(spec["system"], spec["nt"], spec["algorithm"])
    try:
        W.load(s, dict(spec, n_active=700))
        s.SolverOption(O.RelaunchSteps, 300)
        s.SolverOption(O.RelaunchLaunches, 3)
        print("noalloc: duffing RK4, 700 of 1000 instances active", flush=True)
        s.solve_noalloc()
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        out = W.read(s)
    finally:
        s.close()
    ref = oracle.run("duffing", W.subset(spec, np.arange(700)))
    part = {k: (v[..., :700] if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[-1] == 1000 else v)
            for k, v in out.items()}
    assert_bitwise(part, ref, keys=("X", "T", "ACC", "nacc", "nrej", "status"), label="noalloc RK4 active 700")
    assert np.all(out["nacc"][700:] == 0)


def test_noalloc_refuses_handles_beyond_the_scratch(gpu):
    s = ProblemSolver("lorenz", 400000)
    try:
        with pytest.raises(MPGOSError, match="bytes of scratch"):
            s.solve_noalloc()
    finally:
        s.close()
