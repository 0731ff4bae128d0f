"""Python mirror of the reference's per-block coupled-systems ProblemSolver
(SourceCodes/CoupledSystems_PerBlock_Interface.cuh:135-248) over the C ABI.

    s = PBProblemSolver("ring_duffing", ns=4096, ups=256, algorithm=RKCK45, spb=1)
    s.SetHost(system, unit, PB.ActualState, 0, x0)      # unit scope
    s.SetHost(system, PB.SystemParameters, 0, B)        # system scope
    s.SetHost(PB.CouplingIndex, 0, 1)                   # global scope
    s.SetHost(0, PB.CouplingMatrix, row, col, c)        # coupling matrix
    s.SynchroniseFromHostToDevice(PB.All); s.Solve(); ...
"""
import ctypes
import enum

import numpy as np

from ._lib import MPGOSError, check, lib
from .enums import Algorithms, Precision


class PB(enum.IntEnum):
    """Per-block ListOfVariables (PB-IF:35-53) plus new outputs."""
    All = 0
    TimeDomain = 1
    ActualTime = 2
    ActualState = 3
    UnitParameters = 4
    SystemParameters = 5
    GlobalParameters = 6
    IntegerGlobalParameters = 7
    UnitAccessories = 8
    IntegerUnitAccessories = 9
    SystemAccessories = 10
    IntegerSystemAccessories = 11
    CouplingMatrix = 12
    CouplingStrength = 13
    CouplingIndex = 14
    DenseOutput = 15
    DenseIndex = 16
    DenseTime = 17
    DenseState = 18
    AcceptedSteps = 19
    RejectedSteps = 20
    Status = 21
    AdjointState = 22             # new: dL/dx (end in, start out) of the discrete adjoint
    UnitParameterGradient = 23    # new: dL/d(unit parameters)
    SystemParameterGradient = 24  # new: dL/d(system parameters)


class PBOption(enum.IntEnum):
    """Per-block ListOfSolverOptions (PB-IF:55-68)."""
    InitialTimeStep = 0
    ActiveSystems = 1
    MaximumTimeStep = 2
    MinimumTimeStep = 3
    TimeStepGrowLimit = 4
    TimeStepShrinkLimit = 5
    RelativeTolerance = 6
    AbsoluteTolerance = 7
    EventTolerance = 8
    EventDirection = 9
    DenseOutputSaveFrequency = 10
    DenseOutputMinimumTimeStep = 11
    SharedGlobalVariables = 12
    SharedCouplingMatrices = 13
    CouplingKernel = 14          # new: 0 automatic, 1 lane per pair, 2 dense contraction (MFMA)
    MfmaUnitsPerLane = 15        # new: 4 (default) or 2 units per lane in the dense-contraction kernel


class _PBDims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("ns", "ups", "ud", "spb", "nc", "cbw", "cci", "nup", "nsp", "ngp", "nigp",
                                          "nua", "niua", "nsa", "nisa", "ne", "ndo", "algorithm", "precision")]


_UNIT = {PB.ActualState, PB.UnitParameters, PB.UnitAccessories, PB.IntegerUnitAccessories}
_SYSTEM = {PB.TimeDomain, PB.SystemParameters, PB.SystemAccessories, PB.IntegerSystemAccessories, PB.DenseTime,
           PB.CouplingStrength}
_GLOBAL = {PB.GlobalParameters, PB.IntegerGlobalParameters, PB.CouplingIndex}
_SYS1 = {PB.DenseIndex, PB.ActualTime, PB.AcceptedSteps, PB.RejectedSteps, PB.Status}
_TWO_ARG = {PBOption.RelativeTolerance, PBOption.AbsoluteTolerance, PBOption.EventTolerance, PBOption.EventDirection}


def _bind():
    L = lib()
    if getattr(L, "_pb_bound", False):
        return L
    H, c_int, c_double = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "mpgos_pb_list_systems": (ctypes.c_char_p, []),
        "mpgos_pb_create": (c_int, [ctypes.c_char_p] + [c_int] * 9 + [ctypes.POINTER(H)]),
        "mpgos_pb_destroy": (c_int, [H]),
        "mpgos_pb_get_dims": (c_int, [H, ctypes.POINTER(_PBDims)]),
        "mpgos_pb_solver_option": (c_int, [H, c_int, c_int, c_double]),
        "mpgos_pb_set_host": (c_int, [H] + [c_int] * 5 + [c_double]),
        "mpgos_pb_get_host": (c_int, [H] + [c_int] * 5 + [ctypes.POINTER(c_double)]),
        "mpgos_pb_host_buffer": (ctypes.c_void_p, [H, c_int, ctypes.POINTER(ctypes.c_size_t)]),
        "mpgos_pb_device_buffer": (ctypes.c_void_p, [H, c_int, ctypes.POINTER(ctypes.c_size_t)]),
        "mpgos_pb_sync_h2d": (c_int, [H, c_int]),
        "mpgos_pb_sync_d2h": (c_int, [H, c_int]),
        "mpgos_pb_solve": (c_int, [H]),
        "mpgos_pb_synchronise_device": (c_int, [H]),
        "mpgos_pb_insert_synchronisation_point": (c_int, [H]),
        "mpgos_pb_synchronise_solver": (c_int, [H]),
        "mpgos_pb_last_solve_ms": (c_int, [H, ctypes.POINTER(c_double)]),
        "mpgos_pb_stream": (ctypes.c_void_p, [H]),
        "mpgos_pb_adjoint_record": (c_int, [H, c_int, c_int, c_int]),
        "mpgos_pb_adjoint_reverse": (c_int, [H]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L._pb_bound = True
    return L


def list_pb_systems():
    return _bind().mpgos_pb_list_systems().decode().split(",")


class PBProblemSolver:
    def __init__(self, system, ns, ups, algorithm=Algorithms.RKCK45, precision=Precision.DOUBLE, spb=1, cbw=0, cci=0,
                 ndo=0, device=0):
        L = _bind()
        self._h = ctypes.c_void_p()
        check(L.mpgos_pb_create(system.encode(), int(algorithm), int(precision), int(ns), int(ups), int(spb), int(cbw),
                                int(cci), int(ndo), int(device), ctypes.byref(self._h)))
        d = _PBDims()
        check(L.mpgos_pb_get_dims(self._h, ctypes.byref(d)))
        self.d = d
        self.system = system
        self._fp = np.float64 if d.precision == Precision.DOUBLE else np.float32

    def close(self):
        if self._h:
            check(_bind().mpgos_pb_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def SolverOption(self, option, *args):
        option = PBOption(option)
        if len(args) == 1 and option not in _TWO_ARG:
            index, value = 0, args[0]
        elif len(args) == 2 and option in _TWO_ARG:
            index, value = args
        else:
            raise MPGOSError(f"SolverOption: wrong number of arguments for {option.name}")
        check(_bind().mpgos_pb_solver_option(self._h, int(option), int(index), float(value)))

    @staticmethod
    def _args(a):
        # (sys, unit, var, serial, v) | (sys, var, serial, v) | (var, serial, v) |
        # (coupling, CouplingMatrix, row, col, v) | (sys, var, v) | (sys, unit, DenseState, comp, k, v)
        if len(a) == 6:
            return a[2], a[0], a[1], a[3], a[4], a[5]
        if len(a) == 5:
            if PB(a[1]) == PB.CouplingMatrix:
                return a[1], a[0], a[2], a[3], 0, a[4]
            return a[2], a[0], a[1], a[3], 0, a[4]
        if len(a) == 4:
            return a[1], a[0], 0, a[2], 0, a[3]
        if len(a) == 3:
            if isinstance(a[0], PB):
                return a[0], 0, 0, a[1], 0, a[2]
            return a[1], a[0], 0, 0, 0, a[2]
        raise MPGOSError("SetHost: unsupported argument configuration")

    def SetHost(self, *a):
        var, a0, a1, a2, a3, v = self._args(a)
        check(_bind().mpgos_pb_set_host(self._h, int(var), int(a0), int(a1), int(a2), int(a3), float(v)))

    def GetHost(self, *a):
        var, a0, a1, a2, a3, _ = self._args(tuple(a) + (0.0,))
        out = ctypes.c_double()
        check(_bind().mpgos_pb_get_host(self._h, int(var), int(a0), int(a1), int(a2), int(a3), ctypes.byref(out)))
        return out.value

    def host_array(self, var):
        var = PB(var)
        d, fp = self.d, self._fp
        nu = d.ns * d.ups
        table = {
            PB.ActualState: (fp, (d.ud, nu)), PB.UnitParameters: (fp, (d.nup, nu)), PB.UnitAccessories: (fp, (d.nua, nu)),
            PB.IntegerUnitAccessories: (np.int32, (d.niua, nu)), PB.TimeDomain: (fp, (2, d.ns)), PB.ActualTime: (fp, (d.ns,)),
            PB.SystemParameters: (fp, (d.nsp, d.ns)), PB.SystemAccessories: (fp, (d.nsa, d.ns)),
            PB.IntegerSystemAccessories: (np.int32, (d.nisa, d.ns)), PB.CouplingStrength: (fp, (d.nc, d.ns)),
            PB.GlobalParameters: (fp, (d.ngp,)), PB.IntegerGlobalParameters: (np.int32, (d.nigp,)),
            PB.CouplingIndex: (np.int32, (d.nc,)), PB.DenseIndex: (np.int32, (d.ns,)), PB.DenseTime: (fp, (d.ndo, d.ns)),
            PB.DenseState: (fp, (d.ndo, d.ud, nu)), PB.AcceptedSteps: (np.int64, (d.ns,)),
            PB.RejectedSteps: (np.int64, (d.ns,)), PB.Status: (np.int32, (d.ns,)),
            PB.AdjointState: (np.float64, (d.ud, nu)), PB.UnitParameterGradient: (np.float64, (d.nup, nu)),
            PB.SystemParameterGradient: (np.float64, (d.nsp, d.ns)),
        }
        if var == PB.CouplingMatrix:
            n = ctypes.c_size_t()
            p = _bind().mpgos_pb_host_buffer(self._h, int(var), ctypes.byref(n))
            count = n.value // np.dtype(fp).itemsize
            return np.frombuffer((ctypes.c_char * n.value).from_address(p), dtype=fp, count=count) if count else np.zeros(0, fp)
        dtype, shape = table[var]
        count = int(np.prod(shape))
        if count == 0:
            return np.zeros(shape, dtype)
        n = ctypes.c_size_t()
        p = _bind().mpgos_pb_host_buffer(self._h, int(var), ctypes.byref(n))
        if not p:
            raise MPGOSError(lib().mpgos_last_error().decode())
        return np.frombuffer((ctypes.c_char * n.value).from_address(p), dtype=dtype, count=count).reshape(shape)

    def SynchroniseFromHostToDevice(self, var):
        check(_bind().mpgos_pb_sync_h2d(self._h, int(var)))

    def SynchroniseFromDeviceToHost(self, var):
        check(_bind().mpgos_pb_sync_d2h(self._h, int(var)))

    def Solve(self):
        check(_bind().mpgos_pb_solve(self._h))

    def SynchroniseDevice(self):
        check(_bind().mpgos_pb_synchronise_device(self._h))

    def InsertSynchronisationPoint(self):
        check(_bind().mpgos_pb_insert_synchronisation_point(self._h))

    def SynchroniseSolver(self):
        check(_bind().mpgos_pb_synchronise_solver(self._h))

    def last_solve_ms(self):
        ms = ctypes.c_double()
        check(_bind().mpgos_pb_last_solve_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def stream(self):
        return _bind().mpgos_pb_stream(self._h)

    def device_ptr(self, var):
        n = ctypes.c_size_t()
        p = _bind().mpgos_pb_device_buffer(self._h, int(PB(var)), ctypes.byref(n))
        return p, n.value

    # ---------------------------------------------------------------- discrete adjoint
    def adjoint_record(self, checkpoint_interval, checkpoint_capacity, step_capacity):
        """Solve() that also keeps the unit states every `checkpoint_interval`
        accepted steps and (t, h) of every accepted step."""
        check(_bind().mpgos_pb_adjoint_record(self._h, int(checkpoint_interval), int(checkpoint_capacity),
                                              int(step_capacity)))

    def adjoint_reverse(self):
        """Reverse sweep: AdjointState holds dL/dx(end) on entry (device buffer),
        dL/dx(start) on exit; the parameter gradients are filled."""
        check(_bind().mpgos_pb_adjoint_reverse(self._h))

    def adjoint(self, lam_end, checkpoint_interval, checkpoint_capacity, step_capacity):
        """Record, upload dL/dx(end) = lam_end (UD x NS*UPS), reverse, read back.
        Returns (dL/dx(start), dL/d(unit parameters), dL/d(system parameters))."""
        self.adjoint_record(checkpoint_interval, checkpoint_capacity, step_capacity)
        self.host_array(PB.AdjointState)[...] = lam_end
        self.SynchroniseFromHostToDevice(PB.AdjointState)
        self.adjoint_reverse()
        for v in (PB.AdjointState, PB.UnitParameterGradient, PB.SystemParameterGradient):
            self.SynchroniseFromDeviceToHost(v)
        self.InsertSynchronisationPoint()
        self.SynchroniseSolver()
        return tuple(self.host_array(v).copy() for v in (PB.AdjointState, PB.UnitParameterGradient,
                                                          PB.SystemParameterGradient))

    def solve_blocking(self):
        self.Solve()
        self.SynchroniseFromDeviceToHost(PB.All)
        self.InsertSynchronisationPoint()
        self.SynchroniseSolver()
