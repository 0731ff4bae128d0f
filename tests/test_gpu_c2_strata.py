"""GPU parity of the uneven-time-scale workload over all of its strata.

The workload (x'' + 0.2 x' + x + x^3 = B cos(eps t), eps = 10^U[-3, 0],
horizons of 16..4096 forcing periods, RKCK45 at 1e-9, event x' = 0 with 256
logged hits) is sampled at every 512th instance, so every decade of eps is
present, and each horizon is capped at 1e5 time units so that the oracle
finishes in seconds and the longest instance needs about 3e6 attempted steps.
Many instances then cross x' = 0 more often than the event log holds: the
log keeps the first 256 hits while the count goes on (the detection path of
SingleSystem_PerThread_Solver.cuh:323-345 with the hit record of the event
log).  Every kernel form (plain, persistent, relaunch, sorted, host-sliced;
persistent and sorted also replayed from a captured hipGraph) must give the
oracle's bits: states, times, accepted
and rejected step counts, status, event counts and the logged hits.
"""
import numpy as np
import pytest

from helpers import assert_bitwise, run_gpu
from mpgos import workloads as W

pytestmark = pytest.mark.gpu

CAP_T = 1.0e5
EVCAP = 256
KEYS = ("X", "T", "nacc", "nrej", "status", "nev", "evlog", "evlogidx")


@pytest.fixture(scope="module")
def strata(oracle):
    full = W.config2_duffing_uneven(nt=131072, evcap=EVCAP)
    spec = W.subset(full, np.arange(0, 131072, 512))
    spec["TD"] = spec["TD"].copy()
    spec["TD"][1] = np.minimum(spec["TD"][1], CAP_T)
    ref = oracle.run("duffing_eps", spec)
    decades = np.floor(np.log10(spec["cPAR"][2])).astype(int)
    # the sample spans every decade of eps, and the log overflows in several of them
    assert set(decades.tolist()) == {-3, -2, -1}
    over = ref["nev"] > EVCAP
    assert over.sum() > 50 and len(set(decades[over].tolist())) == 3
    # what the log holds at an overflowing instance: 256 hits in time order,
    # each with |x'| below the event tolerance
    i = int(np.argmax(ref["nev"]))
    t = ref["evlog"][:, 0, i]
    assert np.all(np.diff(t) > 0) and np.all(np.abs(ref["evlog"][:, 2, i]) < spec["etol"])
    return spec, ref


@pytest.mark.parametrize("form", ["plain", "persistent", "relaunch", "sorted", "sliced", "persistent-graph", "sorted-graph"])
def test_c2_strata_bitwise(gpu, strata, form):
    spec, ref = strata
    print(f"C2 strata: {spec['nt']} instances, max {int(ref['nacc'].max())} accepted steps, "
          f"max {int(ref['nev'].max())} hits, form {form}", flush=True)
    if form == "plain":
        out = run_gpu(spec)
    elif form == "persistent":
        out = run_gpu(spec, kernel_mode=1)
    elif form == "relaunch":
        out = run_gpu(spec, kernel_mode=2, relaunch_steps=20000, relaunch_launches=40)
    elif form == "sorted":
        out = run_gpu(spec, kernel_mode=3, relaunch_steps=2000)
    elif form == "persistent-graph":
        out = run_gpu(spec, kernel_mode=1, graph=1)
    elif form == "sorted-graph":        # pilot, device sort and final launch replayed from one hipGraph
        out = run_gpu(spec, kernel_mode=3, relaunch_steps=2000, graph=1)
    else:
        out = run_gpu(spec, slice_steps=400000)
        assert out["_slices"] > 3
    assert int(out["nev"].max()) > EVCAP
    assert_bitwise(out, ref, keys=KEYS, label=f"C2 strata {form}")
