"""The drop-in C++ header: a driver written against the reference's
ProblemSolver interface (tests/cpp/lorenz_dropin.cu, the shape of
Tutorials_SingleSystem_PerThread/T2_Lorenz/Lorenz.cu) compiles for gfx950."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_style_driver_compiles(tmp_path):
    pkg = os.path.join(ROOT, "massively-parallel-gpu-ode-solver_amd")
    out = tmp_path / "drv"
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(pkg, "csrc"), "-I" + os.path.join(pkg, "cpp"),
           os.path.join(ROOT, "tests", "cpp", "lorenz_dropin.cu"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert out.exists()
