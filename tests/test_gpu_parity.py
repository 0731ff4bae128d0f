"""GPU parity: the gfx950 kernels against the FP64 oracle on the same seeded
inputs.  FP64 results must be bit-identical (states, times, accessories,
accepted/rejected step counts, event counts, status), which is stronger than
the 1e-12 relative tolerance the target asks for."""
import numpy as np
import pytest

from helpers import assert_bitwise, run_gpu
from mpgos import Algorithms
from mpgos import workloads as W

pytestmark = pytest.mark.gpu


def test_config0_rk4_duffing_small(gpu, oracle):
    spec = W.config0_duffing_rk4(nt=1024)
    ref = oracle.run("duffing", spec)
    out = run_gpu(spec)
    assert (out["nacc"] == 64001).all()
    assert_bitwise(out, ref, label="C0")


def test_config1_lorenz_small(gpu, oracle):
    spec = W.config1_lorenz(nt=4096)
    ref = oracle.run("lorenz", spec)
    out = run_gpu(spec)
    assert_bitwise(out, ref, label="C1 plain")


def test_config1_lorenz_persistent(gpu, oracle):
    spec = W.config1_lorenz(nt=4096)
    ref = oracle.run("lorenz", spec)
    out = run_gpu(spec, kernel_mode=1)
    assert_bitwise(out, ref, label="C1 persistent")


def test_config2_events_subset(gpu, oracle):
    full = W.config2_duffing_uneven(nt=131072, evcap=256)
    # instances with short horizons (the oracle finishes in seconds)
    idx = np.argsort(full["TD"][1])[:512]
    spec = W.subset(full, idx)
    ref = oracle.run("duffing_eps", spec)
    out = run_gpu(spec)
    assert_bitwise(out, ref, keys=("X", "T", "nacc", "nrej", "status", "nev", "evlog", "evlogidx"), label="C2")
    assert ref["nev"].sum() > 0


@pytest.mark.parametrize("mode", [0, 1], ids=["plain", "persistent"])
def test_time_sliced_solve_is_bitwise(gpu, oracle, mode):
    """Solve in slices of a few attempted steps (lane states suspended to HBM
    and resumed): bit-identical to the oracle, events and logs included."""
    full = W.config2_duffing_uneven(nt=131072, evcap=256)
    idx = np.argsort(full["TD"][1])[:700]
    spec = W.subset(full, idx)
    spec["n_active"] = 650
    print(f"sliced solve mode={mode}", flush=True)
    ref = oracle.run("duffing_eps", spec)
    out = run_gpu(spec, kernel_mode=mode, slice_steps=997)
    assert out["_slices"] > 3
    assert_bitwise(out, ref, keys=("X", "T", "nacc", "nrej", "status", "nev", "evlog", "evlogidx"), label="sliced")


@pytest.mark.parametrize("name", ["dense_output_f3_m0.0", "tutorial_duffing_doublewell", "event_ramp_both",
                                  "tutorial_relief_valve"])
def test_time_sliced_scenarios(gpu, oracle, name):
    _, system, spec = next(x for x in S.all_scenarios() if x[0] == name)
    ref = oracle.run(system, spec)
    for mode in (0, 1):
        out = run_gpu(spec, kernel_mode=mode, slice_steps=3)
        keys = ("X", "T", "TD", "cPAR", "ACC", "ACCi", "nacc", "nrej", "status", "nev", "DOidx", "DOt", "DOx")
        assert_bitwise(out, ref, keys=keys, label=f"{name} sliced mode {mode}")


# ---------------------------------------------------------------- scenario sweep
import scenarios as S  # noqa: E402

_ALL = S.all_scenarios()


@pytest.mark.parametrize("mode", [0, 1, 2, 3], ids=["plain", "persistent", "relaunch", "sorted"])
@pytest.mark.parametrize("name,system,spec", _ALL, ids=[x[0] for x in _ALL])
def test_scenarios_bitwise(gpu, oracle, name, system, spec, mode):
    ref = oracle.run(system, spec)
    out = run_gpu(spec, kernel_mode=mode, relaunch_steps=7 if mode in (2, 3) else 0,
                  relaunch_launches=3 if mode == 2 else None)
    keys = ("X", "T", "TD", "cPAR", "ACC", "ACCi", "nacc", "nrej", "status", "nev", "DOidx", "DOt", "DOx",
            "evlog", "evlogidx")
    assert_bitwise(out, ref, keys=keys, label=f"{name}/{mode}")


@pytest.mark.parametrize("block,variant", [(64, 1), (128, 2), (256, 3), (32, 0), (96, 0)])
def test_launch_shapes_do_not_change_results(gpu, oracle, block, variant):
    spec = W.config1_lorenz(nt=3000)
    ref = oracle.run("lorenz", spec)
    for mode in (0, 1, 2, 3):
        out = run_gpu(spec, kernel_mode=mode, block=block, variant=variant)
        assert_bitwise(out, ref, label=f"lorenz block={block} variant={variant} mode={mode}")


def test_persistent_with_fewer_lanes_than_instances(gpu, oracle):
    # one wave per SIMD: lanes are refilled from the queue many times
    spec = W.config1_lorenz(nt=70000)
    ref = oracle.run("lorenz", spec)
    out = run_gpu(spec, kernel_mode=1, waves_per_simd=1)
    assert_bitwise(out, ref, label="lorenz persistent refill")


@pytest.mark.parametrize("solo", [0, 40], ids=["shared", "solo40"])
@pytest.mark.parametrize("slice_steps", [0, 1500], ids=["whole", "sliced"])
def test_persistent_longest_first_queue(gpu, oracle, slice_steps, solo):
    """Work queue in longest-integration-interval-first order (uneven
    horizons, one wave per SIMD so lanes refill, long rounds so compaction is
    often skipped), optionally with the 40 longest instances alone in their
    workgroups: results bit-identical to the oracle, whole and time-sliced."""
    full = W.config2_duffing_uneven(nt=131072, evcap=256)
    idx = np.argsort(full["TD"][1])[:1200]
    spec = W.subset(full, idx[::-1].copy())             # longest horizons last in index order
    print(f"longest-first queue slice_steps={slice_steps} solo={solo}", flush=True)
    ref = oracle.run("duffing_eps", spec)
    out = run_gpu(spec, kernel_mode=1, block=64, waves_per_simd=1, queue_order=1, round_steps=256,
                  slice_steps=slice_steps, solo=solo)
    assert_bitwise(out, ref, keys=("X", "T", "nacc", "nrej", "status", "nev", "evlog", "evlogidx"), label="longest first")


@pytest.mark.parametrize("steps,launches", [(997, 40), (500, 2), (1, 0)], ids=["all_budgeted", "final_unbudgeted", "one_launch"])
def test_relaunch_form_is_bitwise(gpu, oracle, steps, launches):
    """Relaunch form: budgeted launches with the unfinished instances repacked
    (list length on the device), then one unbudgeted launch; events, logs and
    a ragged active count included, repeated Solves continue."""
    full = W.config2_duffing_uneven(nt=131072, evcap=256)
    idx = np.argsort(full["TD"][1])[:700]
    spec = W.subset(full, idx)
    spec["n_active"] = 650
    print(f"relaunch steps={steps} launches={launches}", flush=True)
    ref = oracle.run("duffing_eps", spec)
    outs = run_gpu(spec, kernel_mode=2, relaunch_steps=steps, relaunch_launches=launches, solves=2)
    keys = ("X", "T", "nacc", "nrej", "status", "nev", "evlog", "evlogidx")
    assert_bitwise(outs[0], ref, keys=keys, label="relaunch")
    cur = dict(spec)
    cur["X"], cur["T"] = ref["X"], ref["T"]
    assert_bitwise(outs[1], oracle.run("duffing_eps", cur), keys=keys, label="relaunch second solve")


@pytest.mark.parametrize("pilot", [1, 150, 5000], ids=["pilot1", "pilot150", "pilot_beyond_all"])
def test_sorted_form_is_bitwise(gpu, oracle, pilot):
    """Sorted form: a pilot launch, the unfinished instances ordered on the
    device by estimated remaining steps, one launch to the end; events, logs
    and a ragged active count included, repeated Solves continue."""
    full = W.config2_duffing_uneven(nt=131072, evcap=256)
    idx = np.argsort(full["TD"][1])[:700]
    spec = W.subset(full, idx)
    spec["n_active"] = 650
    print(f"sorted pilot={pilot}", flush=True)
    ref = oracle.run("duffing_eps", spec)
    outs = run_gpu(spec, kernel_mode=3, relaunch_steps=pilot, solves=2)
    keys = ("X", "T", "nacc", "nrej", "status", "nev", "evlog", "evlogidx")
    assert_bitwise(outs[0], ref, keys=keys, label="sorted")
    cur = dict(spec)
    cur["X"], cur["T"] = ref["X"], ref["T"]
    assert_bitwise(outs[1], oracle.run("duffing_eps", cur), keys=keys, label="sorted second solve")


def test_queue_order_option_is_checked(gpu):
    from mpgos import ListOfSolverOptions as O, MPGOSError, ProblemSolver
    s = ProblemSolver("lorenz", 64, 1)
    try:
        with pytest.raises(MPGOSError):
            s.SolverOption(O.QueueOrder, 2)
    finally:
        s.close()


def test_graph_replays_match_eager(gpu, oracle):
    from mpgos import ListOfSolverOptions as O, ListOfVariables as V, ProblemSolver
    spec = W.config1_lorenz(nt=2048)
    ref = oracle.run("lorenz", spec)
    for mode in (0, 1, 2, 3):
        s = ProblemSolver("lorenz", spec["nt"], Algorithms.RKCK45)
        try:
            W.load(s, spec)
            s.SolverOption(O.KernelMode, mode)
            s.SolverOption(O.RelaunchSteps, 200)         # relaunch and sorted forms: several launches in the graph
            s.SolverOption(O.RelaunchLaunches, 4)
            s.SolverOption(O.GraphCapture, 1)
            for rep in range(3):
                W.load(s, spec)                      # eager uploads between replays
                s.solve_blocking()
                out = W.read(s)
                assert_bitwise(out, ref, label=f"graph replay {rep} mode {mode}")
        finally:
            s.close()


def test_repeated_solves_continue_the_integration(gpu, oracle):
    # the reference tutorials chain Solve() calls on the same object (Reference.cu:100-117)
    spec = W.config0_duffing_rk4(nt=256)
    spec["TD"][1] = 2 * W.PI
    outs = run_gpu(spec, solves=3)
    cur = dict(spec)
    for k, out in enumerate(outs):
        ref = oracle.run("duffing", cur)
        assert_bitwise(out, ref, label=f"solve {k}")
        cur = dict(cur)
        cur["X"], cur["T"], cur["ACC"] = ref["X"], ref["T"], ref["ACC"]


def test_keller_miksis_continuation_solves(gpu, oracle):
    # the quasi-periodic forcing tutorial chains Solve() calls, each from the
    # last state to the next maximum of the radius (QuasiperiodicForcing.cu:134-160)
    spec = W.tutorial_keller_miksis(16, 16)
    for mode in (0, 1):
        print(f"keller-miksis continuation mode={mode}", flush=True)
        outs = run_gpu(spec, kernel_mode=mode, solves=3)
        cur = dict(spec)
        for k, out in enumerate(outs):
            ref = oracle.run("keller_miksis", cur)
            assert_bitwise(out, ref, keys=("X", "T", "ACC", "ACCi", "nacc", "nrej", "status"), label=f"solve {k} mode {mode}")
            cur = dict(cur)
            cur["X"], cur["T"] = ref["X"], ref["T"]
        assert (outs[-1]["ACCi"][0] == 1).all()                # every Solve ended at the next maximum


def test_keller_miksis_max_period_solves(gpu, oracle):
    # the large time scale differences tutorial runs one forcing period per
    # Solve, the time restarting at 0 (KellerMiksis.cu:63-90)
    spec = W.tutorial_keller_miksis_max(64)
    print("keller-miksis periods", flush=True)
    outs = run_gpu(spec, solves=4)
    cur = dict(spec)
    for k, out in enumerate(outs):
        ref = oracle.run("keller_miksis_max", cur)
        assert_bitwise(out, ref, keys=("X", "T", "ACC", "nacc", "nrej", "status"), label=f"period {k}")
        cur = dict(cur)
        cur["X"] = ref["X"]
    assert (outs[-1]["ACC"][0] >= outs[-1]["X"][0]).all()


def test_relief_valve_continuation_solves(gpu, oracle):
    # the impact tutorial chains Solve() calls, each from the last state to the
    # next maximum (ImpactDynamics.cu:90-117); state jumps and user terminations
    spec = W.tutorial_relief_valve(nt=256)
    for mode in (0, 1):
        print(f"relief valve continuation mode={mode}", flush=True)
        outs = run_gpu(spec, kernel_mode=mode, solves=4)
        cur = dict(spec)
        for k, out in enumerate(outs):
            ref = oracle.run("relief_valve", cur)
            assert_bitwise(out, ref, keys=("X", "T", "ACC", "ACCi", "nacc", "nrej", "status"), label=f"solve {k} mode {mode}")
            cur = dict(cur)
            cur["X"], cur["T"] = ref["X"], ref["T"]
        assert (outs[-1]["ACCi"][0] == 1).sum() > 200         # most instances ended at a maximum


SPECIAL = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, 5e-324, 1e-310, 800.0, -800.0, 709.78, -745.0])


def test_device_math_matches_host_bitwise(gpu):
    import ctypes
    from test_math import FN, host_eval
    from mpgos.tuning import hip as hip_runtime
    hip = hip_runtime()
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-3e4, 3e4, 100000), rng.uniform(-5, 5, 100000), 10.0 ** rng.uniform(-300, 300, 1000)])
    y = rng.uniform(0.1, 0.3, x.size)
    nbytes = x.nbytes
    bufs = [ctypes.c_void_p() for _ in range(3)]
    for b in bufs:
        assert hip.hipMalloc(ctypes.byref(b), nbytes) == 0
    try:
        L = gpu.lib()
        for name, fn in FN.items():
            arg = np.abs(x) if name in ("log", "pow", "pow_ctrl", "ctrl_root", "sqrt") else (np.clip(x, -700, 700) if name == "exp" else x)
            if name in ("cos", "sin"):      # large arguments: up to 2^52 and the NaN range beyond
                big = 10.0 ** rng.uniform(4, 17, 2000) * np.where(rng.uniform(size=2000) < 0.5, -1.0, 1.0)
                arg = np.concatenate([arg[:-2000], big])
            if name in ("cos", "sin", "exp", "log", "pow", "pow_ctrl", "ctrl_root", "sqrt"):     # special arguments at the end
                arg = np.concatenate([arg[:-len(SPECIAL)], SPECIAL])
            arg = np.ascontiguousarray(arg)
            assert hip.hipMemcpy(bufs[0], arg.ctypes.data, nbytes, 1) == 0
            yy = (np.arange(arg.size) % 2).astype(np.float64) if name == "ctrl_root" else y   # fifth or fourth root
            assert hip.hipMemcpy(bufs[1], yy.ctypes.data, nbytes, 1) == 0
            assert L.mpgos_math_eval_device(fn, bufs[0], bufs[1], bufs[2], arg.size) == 0
            dev = np.empty_like(arg)
            assert hip.hipMemcpy(dev.ctypes.data, bufs[2], nbytes, 2) == 0
            host = host_eval(name, arg, yy if name in ("pow", "pow_ctrl", "ctrl_root") else None)
            both_nan = np.isnan(dev) & np.isnan(host)
            assert np.array_equal(dev.view(np.uint64)[~both_nan], host.view(np.uint64)[~both_nan]), name
            assert np.array_equal(np.isnan(dev), np.isnan(host)), name
    finally:
        for b in bufs:
            hip.hipFree(b)


@pytest.mark.parametrize("precision", [1, 2], ids=["float", "float_kahan"])
def test_fp32_modes_error_against_fp64_oracle(gpu, oracle, precision):
    # FP32 stepping: error against the FP64 oracle is reported, not zero
    spec = W.config0_duffing_rk4(nt=512)
    spec["TD"][1] = 4 * (2 * W.PI)
    ref = oracle.run("duffing", spec)
    out = run_gpu(spec, precision=precision)
    err = np.abs(out["X"].astype(np.float64) - ref["X"]).max()
    assert np.abs(out["nacc"] - ref["nacc"]).max() <= 1      # the horizon itself is rounded to FP32
    # FP32 RK4 over 4000 steps of large-amplitude forced oscillation: plain FP32
    # accumulates ~1e-2; compensated accumulation keeps the state near FP64
    assert err < (5e-2 if precision == 1 else 5e-4), err
    spec = W.config1_lorenz(nt=512)
    spec["TD"][1] = 1.0
    spec["rtol"] = spec["atol"] = 1e-6
    ref = oracle.run("lorenz", spec)
    out = run_gpu(spec, precision=precision)
    rel = np.abs(out["X"].astype(np.float64) - ref["X"]).max() / np.abs(ref["X"]).max()
    assert rel < 1e-3, rel


def _fmt(v):
    """One value as the reference's Print writes it: width 18, precision 10,
    scientific (SingleSystem_PerThread_Interface.cuh:1091-1099)."""
    return f"{v:18.10e}"


@pytest.mark.parametrize("exe_name", ["lorenz_dropin", "lorenz_dropin_san"], ids=["plain", "host_sanitized"])
def test_dropin_cpp_driver_matches_oracle(gpu, oracle, tmp_path, exe_name):
    """A driver written against the reference's per-thread interface, with a
    dense output of 32 points at least 0.5 apart: end states bit for bit, and
    the text of its Print files (general and DenseOutput for two instances)
    character for character equal to the oracle's values written in the
    reference's format.  The same driver is also run with its host code (the
    facade and the whole host engine) built with AddressSanitizer and
    UndefinedBehaviorSanitizer, which must report nothing."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", exe_name)
    assert os.path.exists(exe), f"build() compiles tests/cpp/{exe_name}"
    nt = 4096
    spec = W.config1_lorenz(nt=nt)
    inp = np.concatenate([spec["cPAR"][0], spec["X"].ravel()]).astype(np.float64)
    (tmp_path / "in.bin").write_bytes(inp.tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    print(f"drop-in driver {exe_name}", flush=True)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    out = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.float64).reshape(5, nt)
    dspec = dict(spec)
    dspec["ndo"] = 32
    dspec["options"] = dict(spec["options"], do_min_dt=0.5, do_freq=1)
    ref = oracle.run("lorenz", dspec)
    assert np.array_equal(out[:3].view(np.uint64), ref["X"].view(np.uint64))
    assert np.array_equal(out[3], ref["T"]) and np.array_equal(out[4].astype(np.int64), ref["nacc"])

    def table(name):
        return (tmp_path / name).read_text().split("\n")[:-1]

    rows = table("ActualStateInSolverObject.txt")
    assert rows == [",".join(_fmt(ref["X"][c, i]) for c in range(3)) for i in range(nt)]
    assert table("TimeDomainInSolverObject.txt") == [f"{_fmt(0.0)},{_fmt(10.0)}"] * nt
    assert table("ControlParametersInSolverObject.txt") == [_fmt(v) for v in spec["cPAR"][0]]
    assert table("ActualTimeInSolverObject.txt") == [_fmt(v) for v in ref["T"]]
    for tid in (0, 1234):
        n = int(ref["DOidx"][tid])
        assert 15 <= n <= 32                      # ~one point per 0.5 time units over [0, 10], plus the end
        expect = ["ControlParameters:", _fmt(spec["cPAR"][0, tid]), "SharedParameters:", "",
                  "IntegerSharedParameters:", "", "Accessories:", "", "IntegerAccessories:", "", "",
                  "Time series:"]
        expect += [_fmt(ref["DOt"][k, tid]) + "," + ",".join(_fmt(ref["DOx"][k, c, tid]) for c in range(3))
                   for k in range(n)]
        assert table(f"DenseOutput_{tid}.txt") == expect, tid
        assert float(expect[-1].split(",")[0]) == ref["T"][tid] == 10.0
