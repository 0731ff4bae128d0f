"""Several solver objects on one device and the host problem-pool loop of the
reference tutorials, with bitwise results against the oracle:
  * double buffering (Tutorials_SingleSystem_PerThread/T7_DoubleBuffering/
    DoubleBuffering.cu:89-144): two objects on their own streams, each
    alternating SynchroniseSolver / SynchroniseFromDeviceToHost / Solve /
    InsertSynchronisationPoint while the other one runs;
  * problem pool (T1_Reference/Reference.cu:82-118): a host pool larger than
    the object, processed batch by batch with ActiveNumberOfThreads for the
    ragged last batch (PT-IF:626-628)."""
import numpy as np
import pytest

from helpers import assert_bitwise
from mpgos import ListOfSolverOptions as O, ListOfVariables as V, ProblemSolver
from mpgos import workloads as W

pytestmark = pytest.mark.gpu


def _uneven(nt, lo):
    # short-horizon Duffing instances with the Poincare event (the T7 workload's shape)
    full = W.config2_duffing_uneven(nt=131072, evcap=8)
    idx = np.argsort(full["TD"][1])[lo:lo + nt]
    return W.subset(full, idx)


def test_two_objects_double_buffering(gpu, oracle):
    specs = [_uneven(512, 0), _uneven(512, 512)]
    objs = [ProblemSolver(s["system"], s["nt"], s["algorithm"]) for s in specs]
    try:
        for o, s in zip(objs, specs):
            W.load(o, s)
            o.Solve()
            o.InsertSynchronisationPoint()
        seen = [[], []]
        for _ in range(3):                  # continuation solves, interleaved
            for k, o in enumerate(objs):
                o.SynchroniseSolver()
                o.SynchroniseFromDeviceToHost(V.All)
                o.InsertSynchronisationPoint()
                o.SynchroniseSolver()       # the copy has landed; read it
                seen[k].append(W.read(o))
                o.Solve()
                o.InsertSynchronisationPoint()
        for o in objs:
            o.SynchroniseSolver()
            o.SynchroniseFromDeviceToHost(V.All)
            o.InsertSynchronisationPoint()
            o.SynchroniseSolver()
        for k, o in enumerate(objs):
            seen[k].append(W.read(o))
    finally:
        for o in objs:
            o.close()
    for k, s in enumerate(specs):
        cur = dict(s)
        for j, out in enumerate(seen[k]):
            ref = oracle.run("duffing_eps", cur)
            assert_bitwise(out, ref, keys=("X", "T", "nacc", "nrej", "status", "nev"), label=f"object {k} solve {j}")
            cur = dict(cur)
            cur["X"], cur["T"] = ref["X"], ref["T"]


def test_problem_pool_with_ragged_last_batch(gpu, oracle):
    pool = _uneven(1000, 0)                 # 1000 problems through an object of 384 lanes
    nt = 384
    o = ProblemSolver(pool["system"], nt, pool["algorithm"])
    results = {k: [] for k in ("X", "T", "nacc", "nrej", "nev")}
    try:
        for first in range(0, pool["nt"], nt):
            n = min(nt, pool["nt"] - first)
            batch = W.subset(pool, np.arange(first, first + n))
            if n < nt:                      # pad the object's inputs, only n active
                pad = W.subset(pool, np.concatenate([np.arange(first, first + n), np.zeros(nt - n, int)]))
                batch = dict(pad, n_active=n)
            W.load(o, batch)
            o.solve_blocking()
            out = W.read(o)
            for k in results:
                v = out[k]
                results[k].append(v[..., :n])
    finally:
        o.close()
    got = {k: np.concatenate(v, axis=-1) for k, v in results.items()}
    ref = oracle.run("duffing_eps", pool)
    assert_bitwise(got, ref, keys=tuple(results), label="pool")
