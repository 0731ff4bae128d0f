"""Python mirror of the reference's per-thread ProblemSolver
(SourceCodes/SingleSystem_PerThread_Interface.cuh:117-208) over the C ABI.

Member names, argument order and meaning follow the reference:
    s = ProblemSolver("lorenz", nt=NT, algorithm=RKCK45)
    s.SolverOption(InitialTimeStep, 1e-3)
    s.SolverOption(RelativeTolerance, 0, 1e-10)
    s.SetHost(tid, ActualState, 0, 10.0)
    s.SynchroniseFromHostToDevice(All)
    s.Solve()
    s.SynchroniseFromDeviceToHost(All)
    s.InsertSynchronisationPoint(); s.SynchroniseSolver()
    x0 = s.GetHost(tid, ActualState, 0)
Errors raise MPGOSError instead of terminating the process.  host_array()
gives a numpy view of a pinned host mirror for bulk fills (the vectorised form
of SetHost/GetHost); device_ptr() the device address for HBM-resident use.
"""
import ctypes

import numpy as np

from ._lib import MPGOSError, _Dims, check, lib
from .enums import Algorithms, ListOfSolverOptions, ListOfVariables, Precision

V = ListOfVariables

_TWO_ARG_OPTIONS = {ListOfSolverOptions.RelativeTolerance, ListOfSolverOptions.AbsoluteTolerance,
                    ListOfSolverOptions.EventTolerance, ListOfSolverOptions.EventDirection}


def list_systems():
    return lib().mpgos_pt_list_systems().decode().split(",")


class ProblemSolver:
    def __init__(self, system, nt, algorithm=Algorithms.RKCK45, precision=Precision.DOUBLE, ndo=0, device=0):
        self._h = ctypes.c_void_p()
        check(lib().mpgos_pt_create(system.encode(), int(algorithm), int(precision), int(nt), int(ndo),
                                    int(device), ctypes.byref(self._h)))
        d = _Dims()
        check(lib().mpgos_pt_get_dims(self._h, ctypes.byref(d)))
        self.system = system
        self.NT, self.SD, self.NCP, self.NSP, self.NISP = d.nt, d.sd, d.ncp, d.nsp, d.nisp
        self.NE, self.NA, self.NIA, self.NDO = d.ne, d.na, d.nia, d.ndo
        self.algorithm = Algorithms(d.algorithm)
        self.precision = Precision(d.precision)
        self.has_vjp = bool(d.has_vjp)
        self.event_log_capacity = 0
        self._fp = np.float64 if self.precision == Precision.DOUBLE else np.float32

    # ---------------------------------------------------------------- lifecycle
    def close(self):
        if self._h:
            check(lib().mpgos_pt_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- options
    def SolverOption(self, option, *args):
        option = ListOfSolverOptions(option)
        if len(args) == 1:
            if option in _TWO_ARG_OPTIONS:
                raise MPGOSError(f"SolverOption: {option.name} needs 2 input arguments")
            index, value = 0, args[0]
        elif len(args) == 2:
            if option not in _TWO_ARG_OPTIONS:
                raise MPGOSError(f"SolverOption: {option.name} needs 1 input argument")
            index, value = args
        else:
            raise MPGOSError("SolverOption takes 2 or 3 arguments")
        check(lib().mpgos_pt_solver_option(self._h, int(option), int(index), float(value)))
        if option == ListOfSolverOptions.EventLogCapacity:
            self.event_log_capacity = int(value)

    # ---------------------------------------------------------------- Set/GetHost
    @staticmethod
    def _decode(args):
        # (problem, var, serial, value) | (problem, var, value) | (var, serial, value) |
        # (problem, var, serial, step, value)  -- PT-IF:181-184
        if len(args) == 4:
            return args[1], args[0], args[2], 0, args[3]
        if len(args) == 3:
            if isinstance(args[0], ListOfVariables):
                return args[0], 0, args[1], 0, args[2]
            return args[1], args[0], 0, 0, args[2]
        if len(args) == 5:
            return args[1], args[0], args[2], args[3], args[4]
        raise MPGOSError("SetHost: unsupported argument configuration")

    def SetHost(self, *args):
        var, problem, serial, step, value = self._decode(args)
        check(lib().mpgos_pt_set_host(self._h, int(var), int(problem), int(serial), int(step), float(value)))

    def GetHost(self, *args):
        # (problem, var, serial) | (problem, var) | (var, serial) | (problem, var, serial, step)
        if len(args) == 3:
            var, problem, serial, step = args[1], args[0], args[2], 0
        elif len(args) == 2:
            if isinstance(args[0], ListOfVariables):
                var, problem, serial, step = args[0], 0, args[1], 0
            else:
                var, problem, serial, step = args[1], args[0], 0, 0
        elif len(args) == 4:
            var, problem, serial, step = args[1], args[0], args[2], args[3]
        else:
            raise MPGOSError("GetHost: unsupported argument configuration")
        out = ctypes.c_double()
        check(lib().mpgos_pt_get_host(self._h, int(var), int(problem), int(serial), int(step), ctypes.byref(out)))
        return out.value

    # ---------------------------------------------------------------- bulk access
    def _dtype_shape(self, var):
        fp, NT = self._fp, self.NT
        table = {
            V.TimeDomain: (fp, (2, NT)), V.ActualState: (fp, (self.SD, NT)), V.ActualTime: (fp, (NT,)),
            V.ControlParameters: (fp, (self.NCP, NT)), V.SharedParameters: (fp, (self.NSP,)),
            V.IntegerSharedParameters: (np.int32, (self.NISP,)), V.Accessories: (fp, (self.NA, NT)),
            V.IntegerAccessories: (np.int32, (self.NIA, NT)), V.DenseIndex: (np.int32, (NT,)),
            V.DenseTime: (fp, (self.NDO, NT)), V.DenseState: (fp, (self.NDO, self.SD, NT)),
            V.AcceptedSteps: (np.int64, (NT,)), V.RejectedSteps: (np.int64, (NT,)),
            V.Status: (np.int32, (NT,)), V.EventCount: (np.int32, (NT,)),
            V.EventLog: (fp, (self.event_log_capacity, self.SD + 1, NT)),
            V.EventLogIndex: (np.int32, (self.event_log_capacity, NT)),
            V.AdjointState: (fp, (self.SD, NT)), V.ParameterGradient: (fp, (self.NCP, NT)),
        }
        if var not in table:
            raise MPGOSError(f"no single buffer for {ListOfVariables(var).name}")
        return table[var]

    def host_array(self, var):
        """numpy view (no copy) of the pinned host mirror of `var`, SoA layout."""
        var = ListOfVariables(var)
        dtype, shape = self._dtype_shape(var)
        n = ctypes.c_size_t()
        p = lib().mpgos_pt_host_buffer(self._h, int(var), ctypes.byref(n))
        count = int(np.prod(shape))
        if count == 0:
            return np.zeros(shape, dtype)
        if not p:
            raise MPGOSError(lib().mpgos_last_error().decode())
        buf = (ctypes.c_char * n.value).from_address(p)
        return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)

    def device_ptr(self, var):
        n = ctypes.c_size_t()
        p = lib().mpgos_pt_device_buffer(self._h, int(ListOfVariables(var)), ctypes.byref(n))
        return p, n.value

    # ---------------------------------------------------------------- transfers, solve
    def SynchroniseFromHostToDevice(self, var):
        check(lib().mpgos_pt_sync_h2d(self._h, int(var)))

    def SynchroniseFromDeviceToHost(self, var):
        check(lib().mpgos_pt_sync_d2h(self._h, int(var)))

    def Solve(self):
        check(lib().mpgos_pt_solve(self._h))

    def solve_slice(self, max_attempted_steps):
        """One slice of a Solve (synchronous): every unfinished instance attempts
        at most `max_attempted_steps` steps; returns the number of instances
        still unfinished (0 completes the Solve, bit-identical to Solve())."""
        left = ctypes.c_int()
        check(lib().mpgos_pt_solve_slice(self._h, int(max_attempted_steps), ctypes.byref(left)))
        return left.value

    def solve_noalloc(self):
        """Solve() in the relaunch form with the library's static per-GPU
        scratch memory: allocates nothing (for graph capture and callers that
        cannot allocate).  Bit-identical to Solve(); calls that overlap on one
        GPU share the scratch memory and race."""
        check(lib().mpgos_pt_solve_noalloc(self._h))

    def SynchroniseDevice(self):
        check(lib().mpgos_pt_synchronise_device(self._h))

    def InsertSynchronisationPoint(self):
        check(lib().mpgos_pt_insert_synchronisation_point(self._h))

    def SynchroniseSolver(self):
        check(lib().mpgos_pt_synchronise_solver(self._h))

    def Print(self, var, thread_id=0, directory=None):
        d = directory.encode() if directory else None
        check(lib().mpgos_pt_print(self._h, int(var), int(thread_id), d))

    def last_solve_ms(self):
        ms = ctypes.c_double()
        check(lib().mpgos_pt_last_solve_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def stream(self):
        return lib().mpgos_pt_stream(self._h)

    # ---------------------------------------------------------------- discrete adjoint
    def adjoint_record(self, checkpoint_interval, checkpoint_capacity):
        """Solve() that also keeps (x, t, proposed dt) every `checkpoint_interval`
        accepted steps (at most `checkpoint_capacity` per instance)."""
        check(lib().mpgos_pt_adjoint_record(self._h, int(checkpoint_interval), int(checkpoint_capacity)))

    def adjoint_reverse(self):
        """Reverse sweep: AdjointState holds dL/dx(T) on entry (device buffer),
        dL/dx(0) on exit; ParameterGradient receives dL/dcPAR."""
        check(lib().mpgos_pt_adjoint_reverse(self._h))

    def adjoint(self, lam_T, checkpoint_interval, checkpoint_capacity):
        """Convenience: record, upload dL/dx(T) = lam_T (SD x NT), reverse, read
        back.  Returns (dL/dx(0), dL/dcPAR) as numpy arrays; the forward result
        of the recorded Solve is left in the solver's buffers."""
        self.adjoint_record(checkpoint_interval, checkpoint_capacity)
        self.host_array(V.AdjointState)[...] = lam_T
        self.SynchroniseFromHostToDevice(V.AdjointState)
        self.adjoint_reverse()
        self.SynchroniseFromDeviceToHost(V.AdjointState)
        self.SynchroniseFromDeviceToHost(V.ParameterGradient)
        self.InsertSynchronisationPoint()
        self.SynchroniseSolver()
        return self.host_array(V.AdjointState).copy(), self.host_array(V.ParameterGradient).copy()

    def solve_blocking(self):
        """Solve, copy everything back and wait (the reference's usual sequence)."""
        self.Solve()
        self.SynchroniseFromDeviceToHost(ListOfVariables.All)
        self.InsertSynchronisationPoint()
        self.SynchroniseSolver()
