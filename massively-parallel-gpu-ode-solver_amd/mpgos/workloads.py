"""Seeded synthetic workloads of the benchmark configurations and helpers that
move a workload description ("spec") into and out of a ProblemSolver.

A spec is a dict: system name, algorithm, sizes (nt, ndo, sd, ncp, nsp, nisp,
ne, na, nia), solver options, tolerances and the SoA input arrays in the
reference's layout (index problem + component*NT).  The same spec feeds the
GPU solver and the host oracle, so parity tests compare like with like.

Random numbers: counter-based SplitMix64 (Steele, Lea, Flood 2014) with
seed 20261202 and one stream per configuration; uniforms take the top 53
bits; normals use Box-Muller.
"""
import numpy as np

from .enums import Algorithms, ListOfSolverOptions as O, ListOfVariables as V

SEED = 20261202
PI = 3.14159265358979323846   # the reference tutorials' literal (Reference.cu:10)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)

SYSTEM_DIMS = {
    #              sd ncp nsp nisp ne na nia
    "duffing":     (2, 2, 0, 0, 0, 1, 0),
    "duffing_q":   (3, 2, 0, 0, 0, 0, 0),
    "duffing_eps": (2, 3, 0, 0, 1, 0, 0),
    "duffing_dw":  (2, 1, 1, 0, 2, 3, 1),
    "lorenz":      (3, 1, 0, 0, 0, 0, 0),
    "affine2":     (2, 6, 1, 0, 1, 0, 1),
}


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64 stream: draw i of stream s is mix(base_s + (i+1)*golden)."""

    def __init__(self, seed=SEED, stream=0):
        with np.errstate(over="ignore"):
            self.base = _mix(np.uint64(seed) ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)))
        self.counter = 0

    def u64(self, n):
        i = np.arange(self.counter + 1, self.counter + n + 1, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            return _mix(self.base + i * _GOLDEN)

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n, sigma=1.0):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))          # 1 - u1 in (0, 1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return sigma * z


def linspace(b, e, n):
    """The tutorials' Linspace (Reference.cu:134-150): x[i] = b + i*inc, x[n-1] = e."""
    x = np.empty(n, np.float64)
    x[0] = b
    if n > 1:
        inc = (e - b) / (n - 1)
        x[1:n - 1] = b + np.arange(1, n - 1, dtype=np.float64) * inc
        x[n - 1] = e
    return x


def new_spec(system, algorithm, nt, ndo=0):
    sd, ncp, nsp, nisp, ne, na, nia = SYSTEM_DIMS[system]
    return {
        "system": system, "algorithm": int(algorithm), "nt": nt, "ndo": ndo,
        "sd": sd, "ncp": ncp, "nsp": nsp, "nisp": nisp, "ne": ne, "na": na, "nia": nia,
        "options": {}, "rtol": 1e-8, "atol": 1e-8, "etol": 1e-6, "edir": 0, "evcap": 0,
        "TD": np.zeros((2, nt)), "X": np.zeros((sd, nt)), "T": np.zeros(nt), "cPAR": np.zeros((ncp, nt)),
        "sPAR": np.zeros(nsp), "sPARi": np.zeros(nisp, np.int32), "ACC": np.zeros((na, nt)),
        "ACCi": np.zeros((nia, nt), np.int32), "DOidx": np.zeros(nt, np.int32),
    }


def subset(spec, idx):
    """A spec restricted to the instances `idx` (for bounded CPU samples)."""
    idx = np.asarray(idx)
    s = dict(spec)
    s["nt"] = len(idx)
    s["options"] = dict(spec["options"])
    for k in ("TD", "X", "cPAR", "ACC", "ACCi"):
        s[k] = np.ascontiguousarray(spec[k][:, idx])
    for k in ("T", "DOidx"):
        s[k] = np.ascontiguousarray(spec[k][idx])
    s.pop("n_active", None)
    return s


# ---------------------------------------------------------------- configurations
def config0_duffing_rk4(nt=16384):
    """Fixed-step RK4, hardening Duffing, k = 0.2, B = linspace(0.3, 10),
    (x0, v0) = (-0.5, 0.1), dt = 2 pi/1000 over 64 forcing periods in one Solve
    (64001 accepted steps: the last one is the clip to 128 pi)."""
    s = new_spec("duffing", Algorithms.RK4, nt)
    s["TD"][0] = 0.0
    s["TD"][1] = 64.0 * (2.0 * PI)
    s["X"][0] = -0.5
    s["X"][1] = 0.1
    s["cPAR"][0] = 0.2
    s["cPAR"][1] = linspace(0.3, 10.0, nt)
    s["options"]["dt0"] = 2.0 * PI / 1000.0
    return s


def config1_lorenz(nt=262144):
    """Lorenz (sigma 10, b 8/3), rho ~ U[0, 21], x0 = (10,10,10) + N(0, 0.01)
    (variance 0.01), RKCK45 at rtol = atol = 1e-10 on t in [0, 10]."""
    s = new_spec("lorenz", Algorithms.RKCK45, nt)
    rs = Stream(SEED, 1)
    s["cPAR"][0] = rs.uniform(nt, 0.0, 21.0)
    s["X"][:] = 10.0 + rs.normal(3 * nt, 0.1).reshape(3, nt)
    s["TD"][1] = 10.0
    s["rtol"] = 1e-10
    s["atol"] = 1e-10
    return s


def config2_duffing_uneven(nt=131072, evcap=256):
    """Uneven time scales: x'' + 0.2 x' + x + x^3 = B cos(eps t),
    eps = 10^U[-3, 0], B = linspace(0.3, 10), horizon P 2pi/eps with
    P = round(2^U[4, 12]); RKCK45 at rtol = atol = 1e-9; event x' = 0
    (tolerance 1e-6, both directions), up to `evcap` hits logged."""
    s = new_spec("duffing_eps", Algorithms.RKCK45, nt)
    rs = Stream(SEED, 2)
    eps = 10.0 ** rs.uniform(nt, -3.0, 0.0)
    periods = np.round(2.0 ** rs.uniform(nt, 4.0, 12.0))
    s["cPAR"][0] = 0.2
    s["cPAR"][1] = linspace(0.3, 10.0, nt)
    s["cPAR"][2] = eps
    s["TD"][1] = periods * (2.0 * PI) / eps
    s["X"][0] = -0.5
    s["X"][1] = 0.1
    s["rtol"] = 1e-9
    s["atol"] = 1e-9
    s["etol"] = 1e-6
    s["edir"] = 0
    s["evcap"] = evcap
    return s


def config3_duffing_adjoint(nt=65536, algorithm=Algorithms.RK4, periods=64):
    """Duffing with a quadrature state, (k, B) = (0.2, linspace(0.3, 10)) with
    seeded +-5% perturbations, (x0, v0) = (-0.5, 0.1); RK4 at dt = 2pi/500 or
    RKCK45 at rtol = atol = 1e-9.  The loss of instance i is the mean of x^2
    over the last forcing period, (q(H) - q(H - 2pi)) / 2pi, H = 64 periods.
    Returns the spec of the first segment [0, H - 2pi]; the second segment
    starts from its end state."""
    s = new_spec("duffing_q", algorithm, nt)
    rs = Stream(SEED, 3)
    s["cPAR"][0] = 0.2 * (1.0 + rs.uniform(nt, -0.05, 0.05))
    s["cPAR"][1] = linspace(0.3, 10.0, nt) * (1.0 + rs.uniform(nt, -0.05, 0.05))
    s["X"][0] = -0.5
    s["X"][1] = 0.1
    s["X"][2] = 0.0
    s["TD"][0] = 0.0
    s["TD"][1] = (periods - 1) * (2.0 * PI)
    s["periods"] = periods
    if algorithm == Algorithms.RK4:
        s["options"]["dt0"] = 2.0 * PI / 500.0
    else:
        s["rtol"] = 1e-9
        s["atol"] = 1e-9
    return s


# ---------------------------------------------------------------- spec <-> solver
_OPT = {"dt0": O.InitialTimeStep, "dtmax": O.MaximumTimeStep, "dtmin": O.MinimumTimeStep,
        "grow": O.TimeStepGrowLimit, "shrink": O.TimeStepShrinkLimit,
        "do_min_dt": O.DenseOutputMinimumTimeStep, "do_freq": O.DenseOutputSaveFrequency}


def load(solver, spec):
    """Options, tolerances and inputs of `spec` into `solver`, then H2D (All)."""
    for k, v in spec["options"].items():
        solver.SolverOption(_OPT[k], v)
    if "n_active" in spec:
        solver.SolverOption(O.ActiveNumberOfThreads, spec["n_active"])
    rtol = np.broadcast_to(np.asarray(spec["rtol"], np.float64), (solver.SD,))
    atol = np.broadcast_to(np.asarray(spec["atol"], np.float64), (solver.SD,))
    for i in range(solver.SD):
        solver.SolverOption(O.RelativeTolerance, i, rtol[i])
        solver.SolverOption(O.AbsoluteTolerance, i, atol[i])
    if solver.NE:
        etol = np.broadcast_to(np.asarray(spec["etol"], np.float64), (solver.NE,))
        edir = np.broadcast_to(np.asarray(spec["edir"], np.int64), (solver.NE,))
        for i in range(solver.NE):
            solver.SolverOption(O.EventTolerance, i, etol[i])
            solver.SolverOption(O.EventDirection, i, int(edir[i]))
    if spec.get("evcap", 0):
        solver.SolverOption(O.EventLogCapacity, spec["evcap"])
    fill = {V.TimeDomain: "TD", V.ActualState: "X", V.ActualTime: "T", V.ControlParameters: "cPAR",
            V.SharedParameters: "sPAR", V.IntegerSharedParameters: "sPARi", V.Accessories: "ACC",
            V.IntegerAccessories: "ACCi", V.DenseIndex: "DOidx"}
    for var, key in fill.items():
        dst = solver.host_array(var)
        if dst.size:
            dst[...] = np.asarray(spec[key]).reshape(dst.shape)
    solver.SynchroniseFromHostToDevice(V.All)


def read(solver):
    """Outputs of the last Solve (after SynchroniseFromDeviceToHost(All) and a
    wait) as a dict with the oracle's keys."""
    out = {}
    for var, key in {V.TimeDomain: "TD", V.ActualState: "X", V.ActualTime: "T", V.ControlParameters: "cPAR",
                     V.Accessories: "ACC", V.IntegerAccessories: "ACCi", V.DenseIndex: "DOidx",
                     V.AcceptedSteps: "nacc", V.RejectedSteps: "nrej", V.Status: "status",
                     V.EventCount: "nev"}.items():
        out[key] = solver.host_array(var).copy()
    if solver.NDO:
        out["DOt"] = solver.host_array(V.DenseTime).copy()
        out["DOx"] = solver.host_array(V.DenseState).copy()
    if solver.event_log_capacity:
        solver.SynchroniseFromDeviceToHost(V.EventLog)
        solver.SynchroniseFromDeviceToHost(V.EventLogIndex)
        solver.InsertSynchronisationPoint()
        solver.SynchroniseSolver()
        out["evlog"] = solver.host_array(V.EventLog).copy()
        out["evlogidx"] = solver.host_array(V.EventLogIndex).copy()
    return out
