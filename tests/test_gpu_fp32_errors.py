"""FP32 stepping modes against the FP64 oracle, one workload per configuration
(samples of the benchmark configurations, sizes kept small).  Prints one
table row per configuration and mode: median and maximum relative state error
(|x32 - x64| / max(|x64|, 1e-3)) and the fraction of instances whose
accepted-step count differs from FP64.  FLOAT_KAHAN keeps state and time in
compensated FP32 with the FP64 error norm, controller and event step control
(per-thread and per-block modules).

The GPU results are deterministic, so each configuration and mode is held to
about twice its measured median and maximum error (the bounds below), and the
fixed-step RK4 runs in FLOAT_KAHAN mode must keep the FP64 step counts."""
import numpy as np
import pytest

from helpers import run_gpu
from mpgos import Algorithms, ListOfVariables as V, Precision, ProblemSolver
from mpgos import workloads as W

pytestmark = pytest.mark.gpu

ROWS = []

# (median, max) bounds: about twice the measured values on the MI355X.  For
# plain FP32 on the chaotic Lorenz ensemble the maximum is the distance between
# two unrelated points of the attractor over the size of the FP64 state (it
# moved from 1.8 to 7.0 when the controller's R^(1/5) changed in its last
# bits), so that bound is the attractor's scale, not twice a measured value;
# the FP32 maxima at C2's 1e-9 moved with it (0.014 to 0.044).
BOUNDS = {
    ("C0", "FLOAT"): (2.5e-3, 3.6), ("C0", "FLOAT_KAHAN"): (2e-5, 0.22),
    ("C1", "FLOAT"): (1e-6, 50.0), ("C1", "FLOAT_KAHAN"): (7e-8, 0.19),
    ("C2", "FLOAT"): (2e-5, 0.075), ("C2", "FLOAT_KAHAN"): (1e-5, 0.07),
    ("C3rk4", "FLOAT"): (7e-4, 126.0), ("C3rk4", "FLOAT_KAHAN"): (2.1e-6, 0.094),
    ("C3rkck45", "FLOAT"): (2.3e-3, 92.0), ("C3rkck45", "FLOAT_KAHAN"): (3.9e-6, 0.17),
    ("C4", "FLOAT"): (2.7e-5, 0.14), ("C4", "FLOAT_KAHAN"): (1.3e-6, 4.1e-3),
    ("C2at1e-9", "FLOAT"): (5.7e-5, 9e-2), ("C2at1e-9", "FLOAT_KAHAN"): (3.2e-6, 1.1e-3),
}


def _stats(x32, x64, n32, n64):
    scale = np.maximum(np.abs(x64), 1e-3)
    rel = np.abs(np.asarray(x32, np.float64) - x64) / scale
    return float(np.median(rel)), float(rel.max()), float(np.mean(np.asarray(n32) != np.asarray(n64)))


def _check(key, mode, med, mx, frac, same_steps=False):
    bmed, bmax = BOUNDS[(key, mode)]
    assert med < bmed and mx < bmax, (key, mode, med, mx)
    if same_steps:
        assert frac == 0.0, (key, mode, frac)


def _per_thread(key, name, system, spec, oracle, fixed_step=False):
    ref = oracle.run(system, spec)
    meds = {}
    for prec in (Precision.FLOAT, Precision.FLOAT_KAHAN):
        out = run_gpu(spec, precision=prec)
        med, mx, frac = _stats(out["X"], ref["X"], out["nacc"], ref["nacc"])
        ROWS.append((name, prec.name, med, mx, frac))
        assert np.isfinite(out["X"]).all()
        _check(key, prec.name, med, mx, frac, same_steps=fixed_step and prec == Precision.FLOAT_KAHAN)
        meds[prec] = med
    assert meds[Precision.FLOAT_KAHAN] < meds[Precision.FLOAT]


def test_c0_duffing_rk4(gpu, oracle):
    spec = W.subset(W.config0_duffing_rk4(), np.arange(0, 16384, 16))
    _per_thread("C0", "C0 Duffing RK4 dt 2pi/1000, 64 periods (every 16th of 16384)", "duffing", spec, oracle,
                fixed_step=True)


def test_c1_lorenz(gpu, oracle):
    spec = W.subset(W.config1_lorenz(), np.arange(0, 262144, 64))
    _per_thread("C1", "C1 Lorenz RKCK45 1e-10, t in [0, 10] (every 64th of 262144)", "lorenz", spec, oracle)


def test_c2_uneven(gpu, oracle):
    # compared at 1e-6 over the first 4 forcing periods of each instance; the
    # configuration's own 1e-9 follows below
    full = W.config2_duffing_uneven(evcap=0)
    spec = W.subset(full, np.arange(0, 131072, 512))
    eps = spec["cPAR"][2]
    spec["TD"][1] = np.minimum(spec["TD"][1], 4 * 2 * W.PI / eps)
    spec["rtol"] = spec["atol"] = 1e-6
    _per_thread("C2", "C2 Duffing eps-forced RKCK45 at 1e-6, 4 periods (every 512th of 131072)", "duffing_eps", spec, oracle)


def _sliced_run(spec, prec, budget, max_slices):
    """Time-sliced Solve with a bounded number of slices: returns the outputs
    and whether every instance finished within max_slices * budget attempted
    steps."""
    s = ProblemSolver(spec["system"], spec["nt"], spec["algorithm"], prec)
    try:
        W.load(s, spec)
        left = None
        for _ in range(max_slices):
            left = s.solve_slice(budget)
            if left == 0:
                break
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        return W.read(s), left == 0
    finally:
        s.close()


def test_c2_uneven_at_its_own_tolerance(gpu, oracle):
    # rtol = atol = 1e-9, the configuration's tolerance, over the first 4
    # forcing periods of each instance capped at 2000 time units (FP32 time
    # still resolves the steps there: ulp(2000) = 1.2e-4); time-sliced with a
    # bounded number of slices, so an instance whose step collapses cannot
    # keep the device busy
    full = W.config2_duffing_uneven(evcap=0)
    spec = W.subset(full, np.arange(0, 131072, 512))
    eps = spec["cPAR"][2]
    spec["TD"][1] = np.minimum(np.minimum(spec["TD"][1], 4 * 2 * W.PI / eps), 2000.0)
    ref = oracle.run("duffing_eps", spec)
    for prec in (Precision.FLOAT, Precision.FLOAT_KAHAN):
        out, finished = _sliced_run(spec, prec, 500000, 8)
        assert finished, f"{prec.name}: not finished within 4e6 attempted steps per instance"
        med, mx, frac = _stats(out["X"], ref["X"], out["nacc"], ref["nacc"])
        steps = float(np.sum(out["nacc"] + out["nrej"])) / float(np.sum(ref["nacc"] + ref["nrej"]))
        ROWS.append(("C2 Duffing eps-forced RKCK45 at 1e-9, min(4 periods, 2000) (every 512th of 131072); "
                     f"attempted steps {steps:.2f}x FP64", prec.name, med, mx, frac))
        assert np.isfinite(out["X"]).all()
        _check("C2at1e-9", prec.name, med, mx, frac)


@pytest.mark.parametrize("alg", [Algorithms.RK4, Algorithms.RKCK45], ids=["rk4", "rkck45"])
def test_c3_duffing_q(gpu, oracle, alg):
    spec = W.subset(W.config3_duffing_adjoint(algorithm=alg), np.arange(0, 65536, 64))
    _per_thread(f"C3{alg.name.lower()}", f"C3 Duffing+q {alg.name} forward, 63 periods (every 64th of 65536)", "duffing_q",
                spec, oracle, fixed_step=alg == Algorithms.RK4)


def test_c4_ring(gpu, oracle):
    from mpgos.perblock import PBProblemSolver
    import pyoracle
    spec = W.subset_pb(W.config4_ring(ns=4096, ups=256, periods=4), np.arange(0, 4096, 256))
    ref = pyoracle.run_pb("ring_duffing", spec)
    meds = {}
    for prec in (Precision.FLOAT, Precision.FLOAT_KAHAN):
        s = PBProblemSolver("ring_duffing", spec["ns"], spec["ups"], spec["algorithm"], prec)
        try:
            W.load_pb(s, spec)
            s.solve_blocking()
            out = W.read_pb(s)
        finally:
            s.close()
        med, mx, frac = _stats(out["X"], ref["X"], out["nacc"], ref["nacc"])
        ROWS.append(("C4 ring 256 units RKCK45 1e-9, 4 periods (every 256th of 4096 systems)", prec.name, med, mx, frac))
        assert np.isfinite(out["X"]).all()
        _check("C4", prec.name, med, mx, frac)
        meds[prec] = med
    assert meds[Precision.FLOAT_KAHAN] < meds[Precision.FLOAT]


def test_zz_print_table(gpu):
    print("\n| configuration | mode | median rel. error | max rel. error | step counts differing |")
    print("|---|---|---|---|---|")
    for name, mode, med, mx, frac in ROWS:
        print(f"| {name} | {mode} | {med:.2e} | {mx:.2e} | {100 * frac:.0f} % |")
