"""Exit paths of the drop-in C++ facade (tests/cpp/exit_paths.cu): every
invalid call terminates the process the way the reference does -- a message
on stderr and exit(EXIT_FAILURE) for bad indices and wrong SolverOption
argument counts (SingleSystem_PerThread_Interface.cuh:565-581, 659-663),
exit(1) for a HIP runtime error (the reference's gpuErrCHK, :13-22)."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "exit_paths")

CASES = {
    "set_problem": "ProblemNumber index 256 out of range (acceptable index: 0-255)",
    "set_component": "ActualState index 3 out of range (acceptable index: 0-2)",
    "get_parameter": "ControlParameters index 1 out of range (acceptable index: 0-0)",
    "dense_index": "DenseIndex-Value index 4 out of range (acceptable index: 0-3)",
    "dense_state": "DenseState/TimeStepNumber index 4 out of range (acceptable index: 0-3)",
    "option_needs_index": "This option needs 2 input arguments",
    "option_takes_no_index": "This option needs 1 input arguments",
    "tolerance_index": "AbsoluteTolerance index 3 out of range (acceptable index: 0-2)",
    "print_thread": "Thread index 256 out of range (acceptable index: 0-255)",
}


def _run(mode):
    assert os.path.exists(EXE), "build() compiles tests/cpp/exit_paths"
    return subprocess.run([EXE, mode], capture_output=True, text=True, timeout=120)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the behaviour on a machine without a GPU")
def test_constructor_without_a_device_exits_with_1():
    r = _run("ok")
    assert r.returncode == 1
    assert "In solver member function ProblemSolver" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", sorted(CASES))
def test_invalid_call_exits_with_failure(gpu, mode):
    print(f"facade exit path {mode}", flush=True)
    r = _run(mode)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert CASES[mode] in r.stderr, r.stderr
    assert "returned" not in r.stdout


@pytest.mark.gpu
def test_valid_calls_return(gpu):
    r = _run("ok")
    assert r.returncode == 0 and r.stdout.startswith("ok 1"), (r.returncode, r.stdout, r.stderr)
