"""The oracle's C++ (per-thread and per-block restatements, dual-number
sensitivities, op counting) under AddressSanitizer and
UndefinedBehaviorSanitizer: the oracle pin tests run in a child Python process
against oracle/liboracle_san.so with the sanitizer runtimes preloaded, and must
pass with no sanitizer report.  (The host engine behind the C ABI and the C++
facade runs under the same sanitizers in the drop-in driver GPU test.)"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()


def test_oracle_pins_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=900)
    lib = os.path.join(ROOT, "oracle", "liboracle_san.so")
    env = dict(os.environ, LD_PRELOAD=f"{_runtime('libasan.so')}:{_runtime('libubsan.so')}",
               ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1", MPGOS_ORACLE_LIB=lib)
    probe = subprocess.run([sys.executable, "-c", "import sys; sys.path.insert(0, 'oracle'); import pyoracle; "
                            "pyoracle.lib(); print(pyoracle.LIB)"], cwd=ROOT, env=env, capture_output=True, text=True,
                           timeout=300)
    assert probe.returncode == 0 and probe.stdout.strip() == lib, probe.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py", "tests/test_pb_oracle.py"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=1800)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
