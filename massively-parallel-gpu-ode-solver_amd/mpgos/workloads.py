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
    "relief_valve": (3, 1, 4, 0, 2, 3, 2),
    "keller_miksis": (2, 21, 0, 0, 1, 4, 1),
    "keller_miksis_max": (2, 13, 0, 0, 0, 1, 0),
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


def tutorial_relief_valve(nt=256):
    """The reference's impact-dynamics tutorial setup
    (Tutorials_SingleSystem_PerThread/T6_ImpactDynamics/ImpactDynamics.cu:56-66,
    166-186): pressure relief valve, flow rate q = linspace(0.2, 10) per
    instance, (x, v, p) = (0.2, 0, 0), (kappa, delta, beta, r) =
    (1.25, 10, 20, 0.8), time domain [0, 1e10] (a Solve ends at the next
    maximum of x or when the valve rests), RKCK45 at rtol = atol = 1e-10, both
    event directions -1; repeated Solves continue from the last state."""
    s = new_spec("relief_valve", Algorithms.RKCK45, nt)
    s["TD"][1] = 1.0e10
    s["X"][0] = 0.2
    s["cPAR"][0] = linspace(0.2, 10.0, nt)
    s["sPAR"][:] = (1.25, 10.0, 20.0, 0.8)
    s["rtol"] = 1e-10
    s["atol"] = 1e-10
    s["edir"] = np.array([-1, -1])
    return s


def keller_miksis_parameters(pa1, f1, pa2, f2, phase=0.0, re_um=10.0, pinf_bar=1.0, kappa=1.4):
    """The 21 control parameters of one Keller-Miksis instance from pressure
    amplitudes [bar] and frequencies [kHz], as the reference's quasi-periodic
    forcing tutorial forms them (QuasiperiodicForcing.cu:248-322; water at
    room temperature)."""
    pv, rho, st, vis, cl = 3.166775638952003e+03, 9.970639504998557e+02, 0.071977583160056, \
        8.902125058209557e-04, 1.497251785455527e+03
    pinf, PA1, PA2, RE = pinf_bar * 1.0e5, pa1 * 1.0e5, pa2 * 1.0e5, re_um / 1.0e6
    w1, w2 = 2.0 * PI * (f1 * 1000), 2.0 * PI * (f2 * 1000)
    g = (2.0 * PI / RE / w1) ** 2.0
    return [(2.0 * st / RE + pinf - pv) * g / rho,
            (1.0 - 3.0 * kappa) * (2 * st / RE + pinf - pv) * (2.0 * PI / RE / w1) / cl / rho,
            (pinf - pv) * g / rho,
            (2.0 * st / RE / rho) * g,
            (4.0 * vis / rho / RE ** 2.0) * (2.0 * PI / w1),
            PA1 * g / rho,
            PA2 * g / rho,
            (RE * w1 * PA1 / rho / cl) * g,
            (RE * w2 * PA2 / rho / cl) * g,
            RE * w1 / (2.0 * PI) / cl,
            3.0 * kappa,
            f2 / f1,
            phase,
            2.0 * PI / w1, RE, pa1, f1, pa2, f2, phase, re_um]


def tutorial_keller_miksis(n1=16, n2=16, pa1=0.5, pa2=0.7):
    """The reference's quasi-periodic forcing tutorial setup
    (Tutorials_SingleSystem_PerThread/T4_QuasiperiodicForcing/QuasiperiodicForcing.cu:37-72, 248-322)
    on an n1 x n2 grid of the two frequencies, logspace(20, 1000) kHz each:
    (y1, y2) = (1, 0), time domain [0, 1e10] (a Solve ends at the next local
    maximum of the radius), RKCK45 at rtol = atol = 1e-10, event direction -1;
    repeated Solves continue from the last state."""
    nt = n1 * n2
    s = new_spec("keller_miksis", Algorithms.RKCK45, nt)
    f1 = np.logspace(np.log10(20.0), np.log10(1000.0), n1)
    f2 = np.logspace(np.log10(20.0), np.log10(1000.0), n2)
    for i in range(nt):
        s["cPAR"][:, i] = keller_miksis_parameters(pa1, f1[i % n1], pa2, f2[i // n1])
    s["TD"][1] = 1.0e10
    s["X"][0] = 1.0
    s["rtol"] = 1e-10
    s["atol"] = 1e-10
    s["edir"] = np.array([-1])
    return s


def tutorial_keller_miksis_max(nt=64):
    """The reference's large time scale differences tutorial setup
    (Tutorials_SingleSystem_PerThread/T5_LargeTimeScaleDifferences/KellerMiksis.cu:36-58, 151-210):
    one forcing of 1.5 bar at logspace(20, 1000) kHz per instance, (y1, y2) =
    (1, 0), time domain [0, 1] (one forcing period per Solve, the time
    restarting at 0), RKCK45 at rtol = atol = 1e-10; ACC[0] = largest radius
    of a Solve."""
    s = new_spec("keller_miksis_max", Algorithms.RKCK45, nt)
    f1 = np.logspace(np.log10(20.0), np.log10(1000.0), nt)
    for i in range(nt):
        s["cPAR"][:, i] = keller_miksis_parameters(1.5, f1[i], 0.0, 0.0)[:13]
    s["TD"][1] = 1.0
    s["X"][0] = 1.0
    s["rtol"] = 1e-10
    s["atol"] = 1e-10
    return s


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


# ---------------------------------------------------------------- per-block specs
PB_SYSTEM_DIMS = {
    #               ud nc nup nsp ngp nigp nua niua nsa nisa ne
    "ring_duffing": (2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0),
    "ring_dw":      (2, 2, 3, 0, 1, 0, 3, 1, 2, 1, 1),
    "pb_linear":    (1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0),
    "ring_duffing_q": (3, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0),
}


def coupling_storage_size(ups, cbw, cci):
    if cci == 0 and cbw == 0:
        return ups * ups
    if cci == 0:
        return ups * (2 * cbw + 1)
    return 2 * cbw + 1 if cbw > 0 else ups


def store_coupling(dense, cbw, cci):
    """A full UPS x UPS matrix in the reference's storage form (PB-IF:994-1081):
    returns the stored vector (entries outside the band must be zero, and a
    circular form must be constant along its diagonals)."""
    ups = dense.shape[0]
    if cci == 0 and cbw == 0:
        return np.asarray(dense, np.float64).T.ravel().copy()          # M[row + col*UPS]
    W = 2 * cbw + 1 if cbw > 0 else ups
    out = np.zeros(ups * W if cci == 0 else W)
    for r in range(ups):
        for c in range(ups):
            mc = (c - r + cbw) % ups
            if mc >= W:
                continue
            if cci == 0:
                out[r + mc * ups] = dense[r, c]
            else:
                out[mc] = dense[r, c]
    return out


def new_pb_spec(system, algorithm, ns, ups, cbw=0, cci=0, ndo=0):
    ud, nc, nup, nsp, ngp, nigp, nua, niua, nsa, nisa, ne = PB_SYSTEM_DIMS[system]
    nu = ns * ups
    return {
        "system": system, "algorithm": int(algorithm), "ns": ns, "ups": ups, "cbw": cbw, "cci": cci, "ndo": ndo,
        "ud": ud, "nc": nc, "nup": nup, "nsp": nsp, "ngp": ngp, "nigp": nigp, "nua": nua, "niua": niua, "nsa": nsa,
        "nisa": nisa, "ne": ne, "options": {}, "rtol": 1e-8, "atol": 1e-8, "etol": 1e-6, "edir": 0,
        "X": np.zeros((ud, nu)), "uPAR": np.zeros((nup, nu)), "uACC": np.zeros((nua, nu)),
        "uACCi": np.zeros((niua, nu), np.int32), "TD": np.zeros((2, ns)), "T": np.zeros(ns),
        "sPAR": np.zeros((nsp, ns)), "sACC": np.zeros((nsa, ns)), "sACCi": np.zeros((nisa, ns), np.int32),
        "strength": np.zeros((nc, ns)), "DOidx": np.zeros(ns, np.int32), "gPAR": np.zeros(ngp),
        "gPARi": np.zeros(nigp, np.int32), "cIndex": np.zeros(nc, np.int32),
        "cMat": np.zeros(nc * coupling_storage_size(ups, cbw, cci)),
    }


def config4_ring(ns=4096, ups=256, periods=4, algorithm=Algorithms.RKCK45, system="ring_duffing"):
    """Per-block ring of U forced Duffing units with a dense coupling matrix on
    the velocity equation: C_uv ~ N(0, 0.05/U) (variance), zero diagonal, one
    matrix shared by all systems; per-unit k = 0.2, per-system B ~ U[0.5, 5];
    unit initial states (x, v) ~ U[-1, 1]^2; RKCK45 at rtol = atol = 1e-9 over
    `periods` forcing periods (4 by default)."""
    s = new_pb_spec(system, algorithm, ns, ups)
    rs = Stream(SEED, 4)
    C = rs.normal(ups * ups, np.sqrt(0.05 / ups)).reshape(ups, ups)
    np.fill_diagonal(C, 0.0)
    s["dense_coupling"] = C
    s["cMat"] = store_coupling(C, 0, 0)
    s["uPAR"][0] = 0.2
    s["sPAR"][0] = rs.uniform(ns, 0.5, 5.0)
    s["X"][:2] = rs.uniform(2 * ns * ups, -1.0, 1.0).reshape(2, ns * ups)
    s["TD"][1] = periods * (2.0 * PI)
    s["strength"][0] = 1.0
    s["cIndex"][0] = 1
    s["rtol"] = 1e-9
    s["atol"] = 1e-9
    return s


_PB_OPT = {"dt0": 0, "dtmax": 2, "dtmin": 3, "grow": 4, "shrink": 5, "do_freq": 10, "do_min_dt": 11,
           "shared_globals": 12, "shared_matrix": 13}


def subset_pb(spec, idx):
    """A per-block spec restricted to the systems `idx`."""
    idx = np.asarray(idx)
    ups = spec["ups"]
    units = (idx[:, None] * ups + np.arange(ups)[None, :]).ravel()
    s = dict(spec)
    s["ns"] = len(idx)
    s["options"] = dict(spec["options"])
    for k in ("X", "uPAR", "uACC", "uACCi"):
        s[k] = np.ascontiguousarray(spec[k][:, units])
    for k in ("TD", "sPAR", "sACC", "sACCi", "strength"):
        s[k] = np.ascontiguousarray(spec[k][:, idx])
    for k in ("T", "DOidx"):
        s[k] = np.ascontiguousarray(spec[k][idx])
    s.pop("n_active", None)
    return s


def subset_pb_out(out, idx, ups):
    """Per-block outputs (read_pb) restricted to the systems `idx`."""
    idx = np.asarray(idx)
    units = (idx[:, None] * ups + np.arange(ups)[None, :]).ravel()
    r = {}
    for k, v in out.items():
        if k in ("X", "uACC", "uACCi"):
            r[k] = v[:, units]
        elif k == "DOx":
            r[k] = v[:, :, units]
        elif k in ("TD", "sACC", "sACCi", "DOt", "sPAR", "strength"):
            r[k] = v[:, idx]
        elif k in ("T", "DOidx", "nacc", "nrej", "status"):
            r[k] = v[idx]
    return r


def load_pb(solver, spec):
    """Options and inputs of a per-block spec into a PBProblemSolver, then H2D."""
    from .perblock import PB
    for k, v in spec["options"].items():
        solver.SolverOption(_PB_OPT[k], v)
    if "n_active" in spec:
        solver.SolverOption(1, spec["n_active"])
    ud, ne = spec["ud"], spec["ne"]
    rtol = np.broadcast_to(np.asarray(spec["rtol"], np.float64), (ud,))
    atol = np.broadcast_to(np.asarray(spec["atol"], np.float64), (ud,))
    for i in range(ud):
        solver.SolverOption(6, i, rtol[i])
        solver.SolverOption(7, i, atol[i])
    if ne:
        etol = np.broadcast_to(np.asarray(spec["etol"], np.float64), (ne,))
        edir = np.broadcast_to(np.asarray(spec["edir"], np.int64), (ne,))
        for i in range(ne):
            solver.SolverOption(8, i, etol[i])
            solver.SolverOption(9, i, int(edir[i]))
    fill = {PB.ActualState: "X", PB.UnitParameters: "uPAR", PB.UnitAccessories: "uACC",
            PB.IntegerUnitAccessories: "uACCi", PB.TimeDomain: "TD", PB.ActualTime: "T",
            PB.SystemParameters: "sPAR", PB.SystemAccessories: "sACC", PB.IntegerSystemAccessories: "sACCi",
            PB.CouplingStrength: "strength", PB.DenseIndex: "DOidx", PB.GlobalParameters: "gPAR",
            PB.IntegerGlobalParameters: "gPARi", PB.CouplingIndex: "cIndex", PB.CouplingMatrix: "cMat"}
    for var, key in fill.items():
        dst = solver.host_array(var)
        if dst.size:
            dst[...] = np.asarray(spec[key]).reshape(dst.shape)
    solver.SynchroniseFromHostToDevice(PB.All)


def read_pb(solver):
    from .perblock import PB
    out = {}
    for var, key in {PB.ActualState: "X", PB.UnitAccessories: "uACC", PB.IntegerUnitAccessories: "uACCi",
                     PB.TimeDomain: "TD", PB.ActualTime: "T", PB.SystemAccessories: "sACC",
                     PB.IntegerSystemAccessories: "sACCi", PB.DenseIndex: "DOidx", PB.AcceptedSteps: "nacc",
                     PB.RejectedSteps: "nrej", PB.Status: "status"}.items():
        out[key] = solver.host_array(var).copy()
    if solver.d.ndo:
        out["DOt"] = solver.host_array(PB.DenseTime).copy()
        out["DOx"] = solver.host_array(PB.DenseState).copy()
    return out


def tutorial_ring_dw(algorithm=Algorithms.RKCK45, ns=25, ups=14):
    """The reference's per-block tutorial (T1_RingOfDuffingOscillators/
    RingOfDuffingOscillators.cu:17-270): 25 systems of 14 double-well Duffing
    units on a circular nearest-neighbour ring (CBW 1, CCI 1, two couplings
    of strength 0.004/2), unit 11 started from a 5 x 5 grid of initial
    conditions and every other unit from (-5, -5); uPAR = (0.5, 0.24, 13.633),
    gPAR = 0.004, horizon 10 periods of 2pi/0.5, dt0 = 1e-2, tolerances
    1e-8, event x' = 0 in direction -1 (local maxima), dense output of 500
    points with minimum spacing 1e-6."""
    s = new_pb_spec("ring_dw", algorithm, ns, ups, cbw=1, cci=1, ndo=500)
    nx = ny = int(round(np.sqrt(ns)))
    icx = linspace(-8.0, 8.0, nx)
    icy = linspace(-20.0, 40.0, ny)
    cps = 0.004 / 2.0
    for sid in range(ns):
        x0, y0 = icx[sid // ny], icy[sid % ny]
        s["TD"][1, sid] = 10.0 * 2.0 * PI / 0.5
        s["strength"][:, sid] = cps
        s["sACC"][0, sid] = 1e-2
        s["sACC"][1, sid] = 1e-2
        for u in range(ups):
            g = sid * ups + u
            a, b = (x0, y0) if u == 11 else (-5.0, -5.0)
            s["X"][:, g] = (a, b)
            s["uACC"][0, g] = a
            s["uACC"][1, g] = b
            s["uPAR"][:, g] = (0.5, 0.24, 13.633)
    s["cIndex"][:] = (0, 1)
    s["gPAR"][0] = 0.004
    s["cMat"][:] = np.tile([1.0, 0.0, 1.0], 2)
    s["options"].update(dt0=1e-2, do_min_dt=1e-6, do_freq=1, shared_globals=0)
    s["rtol"] = 1e-8
    s["atol"] = 1e-8
    s["edir"] = -1
    s["spb"] = 3
    return s
