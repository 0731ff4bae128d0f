"""bench.py's output contract without a GPU: more than one GPU is refused with
a message, and every committed measurement line under profiles/ carries the
fields the contract names (metric, value, unit, counts, timing, dtype, data,
the workload config, the roofline object and the CPU baseline object)."""
import glob
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LINES = sorted(glob.glob(os.path.join(ROOT, "profiles", "bench_*.json")))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_there_are_measurement_lines():
    names = {os.path.basename(p) for p in LINES}
    assert {"bench_lorenz_rkck45_262144.json", "bench_ring_duffing_4096x256.json",
            "bench_duffing_uneven_relaunch.json", "bench_duffing_uneven_sorted.json", "bench_adjoint_rk4.json"} <= names


@pytest.mark.parametrize("path", LINES, ids=[os.path.basename(p) for p in LINES])
def test_measurement_line_fields(path):
    with open(path) as f:
        d = json.loads(f.read().strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["value"] > 0 and d["ms_per_step"] > 0
    assert "workload" in d["config"] and "model" not in d["config"]
    r = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in r, k
    # per-thread kernels issue no matrix instructions: FP64 (or FP32) vector ALU bound
    assert r["bound"] in ("hbm", "mfma", "fp64_valu", "fp32_valu") and r["unit"] in ("GB/s", "TFLOP/s")
    if "instances" in d["config"]:
        assert r["bound"] in ("fp64_valu", "fp32_valu"), r["bound"]
    if "systems" in d["config"]:
        assert r["bound"] == "mfma"
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-9
    if r["traffic"] is not None:
        assert r["traffic"] > 0
    c = d["cpu_baseline"]
    if c is not None:
        for k in ("value", "unit", "cores", "kind", "sample"):
            assert k in c, k
        assert c["kind"] in ("port", "reference") and c["cores"] >= 1


def test_default_line_carries_the_whole_metric():
    """The headline line (default run) also carries the adjoint/forward cost of
    both steppers and the per-block ring with its CPU sample and parity."""
    with open(os.path.join(ROOT, "profiles", "bench_lorenz_rkck45_262144.json")) as f:
        d = json.loads(f.read().strip().splitlines()[-1])
    for alg in ("rk4", "rkck45"):
        a = d["adjoint"][alg]
        assert a["adjoint_over_forward"] > 0 and a["checkpoint_interval_from"] == "autotuner"
        assert a["cpu_baseline"]["cores"] >= 1
    r = d["ring"]
    assert r["roofline"]["bound"] == "mfma" and r["parity"]["bitwise_equal_to_oracle"]
    assert r["cpu_baseline"]["value"] > 0 and r["adjoint"]["adjoint_over_forward"] > 0
