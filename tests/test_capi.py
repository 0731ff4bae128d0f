"""The C ABI without a GPU: libmpgos.so loads, exports every entry point that
include/mpgos.h declares, reports errors through return codes instead of
exiting, and the enumerations match the reference's order."""
import ctypes
import os
import re
import subprocess

import pytest

import mpgos
from mpgos import ListOfSolverOptions, ListOfVariables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(mpgos_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", mpgos.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (mpgos_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    L = mpgos.lib()
    for s in syms:
        getattr(L, s)


def test_registered_systems():
    assert set(mpgos.list_systems()) >= {"duffing", "duffing_q", "duffing_eps", "duffing_dw", "lorenz", "affine2",
                                         "relief_valve", "keller_miksis", "keller_miksis_max"}


def test_errors_are_return_codes():
    L = mpgos.lib()
    h = ctypes.c_void_p()
    rc = L.mpgos_pt_create(b"no_such_system", 1, 0, 16, 0, 0, ctypes.byref(h))
    assert rc == -4 and b"unknown system" in L.mpgos_last_error()
    assert L.mpgos_pt_solve(None) == -1
    assert L.mpgos_pt_destroy(None) == -1
    assert L.mpgos_math_eval_host(9, None, None, None, 0) == -1


def test_enumerations_follow_the_reference_order():
    # SingleSystem_PerThread_Interface.cuh:35-62
    ref_vars = ["All", "TimeDomain", "ActualState", "ActualTime", "ControlParameters", "SharedParameters",
                "IntegerSharedParameters", "Accessories", "IntegerAccessories", "DenseOutput", "DenseIndex",
                "DenseTime", "DenseState"]
    ref_opts = ["ThreadsPerBlock", "InitialTimeStep", "ActiveNumberOfThreads", "MaximumTimeStep",
                "MinimumTimeStep", "TimeStepGrowLimit", "TimeStepShrinkLimit", "RelativeTolerance",
                "AbsoluteTolerance", "EventTolerance", "EventDirection", "DenseOutputMinimumTimeStep",
                "DenseOutputSaveFrequency", "PreferSharedMemory"]
    assert [ListOfVariables(i).name for i in range(len(ref_vars))] == ref_vars
    assert [ListOfSolverOptions(i).name for i in range(len(ref_opts))] == ref_opts
    header = open(os.path.join(ROOT, "include", "mpgos.h")).read()
    assert re.search(r"MPGOS_PT_ALL = 0,\s*MPGOS_PT_TIME_DOMAIN,\s*MPGOS_PT_ACTUAL_STATE,", header)
    # options past the reference's: the C enum order equals the Python enum's values
    body = header[header.index("MPGOS_PT_THREADS_PER_BLOCK = 0"):header.index("MPGOS_PT_NUM_OPTIONS")]
    names = re.findall(r"^\s*(MPGOS_PT_[A-Z_]+)", body, re.M)
    assert names.index("MPGOS_PT_PERSISTENT_ROUND_STEPS") == ListOfSolverOptions.PersistentRoundSteps
    assert names.index("MPGOS_PT_QUEUE_ORDER") == ListOfSolverOptions.QueueOrder
    assert names.index("MPGOS_PT_PERSISTENT_SOLO_INSTANCES") == ListOfSolverOptions.PersistentSoloInstances


def test_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    L = mpgos.lib()
    h = ctypes.c_void_p()
    rc = L.mpgos_pt_create(b"lorenz", 1, 0, 16, 0, 0, ctypes.byref(h))
    assert rc == -2 and not h.value


def test_per_block_registry_and_errors():
    assert set(mpgos.list_pb_systems()) >= {"ring_duffing", "ring_dw", "pb_linear"}
    L = mpgos.lib()
    h = ctypes.c_void_p()
    rc = L.mpgos_pb_create(b"no_such_system", 1, 0, 4, 8, 1, 0, 0, 0, 0, ctypes.byref(h))
    assert rc == -4 and b"unknown system" in L.mpgos_last_error()
    assert L.mpgos_pb_solve(None) == -1
    assert L.mpgos_pb_destroy(None) == -1


def test_per_block_enumerations_follow_the_reference_order():
    # CoupledSystems_PerBlock_Interface.cuh:35-68
    from mpgos import PB, PBOption
    ref_vars = ["All", "TimeDomain", "ActualTime", "ActualState", "UnitParameters", "SystemParameters",
                "GlobalParameters", "IntegerGlobalParameters", "UnitAccessories", "IntegerUnitAccessories",
                "SystemAccessories", "IntegerSystemAccessories", "CouplingMatrix", "CouplingStrength",
                "CouplingIndex", "DenseOutput", "DenseIndex", "DenseTime", "DenseState"]
    ref_opts = ["InitialTimeStep", "ActiveSystems", "MaximumTimeStep", "MinimumTimeStep", "TimeStepGrowLimit",
                "TimeStepShrinkLimit", "RelativeTolerance", "AbsoluteTolerance", "EventTolerance",
                "EventDirection", "DenseOutputSaveFrequency", "DenseOutputMinimumTimeStep",
                "SharedGlobalVariables", "SharedCouplingMatrices"]
    assert [PB(i).name for i in range(len(ref_vars))] == ref_vars
    assert [PBOption(i).name for i in range(len(ref_opts))] == ref_opts


def test_per_block_create_without_device_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    L = mpgos.lib()
    h = ctypes.c_void_p()
    rc = L.mpgos_pb_create(b"ring_duffing", 1, 0, 4, 8, 1, 0, 0, 0, 0, ctypes.byref(h))
    assert rc == -2 and not h.value
