"""Pins of the FP64 oracle (oracle/pt_oracle.h) against what mathematics fixes.

The reference ships no golden vectors or known-answer tests (its tutorials are
judged by plots), and its solver cannot be built here (it needs the CUDA
toolchain), so the oracle is pinned by closed forms, invariants, brute-force
time grids and SciPy's DOP853, each chosen so that a plausible slip (a wrong
tableau coefficient, exponent, sign, clip rule or callback order) fails."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import scenarios as S
from mpgos import Algorithms
from mpgos import workloads as W


def run(oracle, spec):
    return oracle.run(spec["system"], spec)


# ------------------------------------------------------------------ steppers
def test_rk4_single_step_is_the_degree4_taylor_polynomial(oracle):
    # RK4 on x' = lam x advances by R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 exactly
    spec = S.affine(nt=1, a00=-1.3, a11=0.7, x0=(1.0, -2.0), td=(0.0, 0.1), dt0=0.2, algorithm=Algorithms.RK4)
    o = run(oracle, spec)
    assert o["nacc"][0] == 1          # dt0 > horizon: one clipped step, then stop
    for lam, x0, got in ((-1.3, 1.0, o["X"][0, 0]), (0.7, -2.0, o["X"][1, 0])):
        z = lam * 0.1
        R = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        assert got == pytest.approx(R * x0, rel=4e-16)


def _fixed_step_duffing(algorithm, h, t1):
    s = W.new_spec("duffing", algorithm, 1)
    s["TD"][1] = t1
    s["X"][0], s["X"][1] = -0.5, 0.1
    s["cPAR"][0], s["cPAR"][1] = 0.2, 2.0
    s["options"].update(dt0=h, grow=1.0, shrink=1.0)
    s["rtol"] = s["atol"] = 1e3
    return s


def _duffing_exact(t1):
    f = lambda t, y: [y[1], 2.0 * math.cos(t) - 0.2 * y[1] - y[0] - y[0] ** 3]
    return solve_ivp(f, (0, t1), [-0.5, 0.1], method="DOP853", rtol=1e-14, atol=1e-14).y[:, -1]


@pytest.mark.parametrize("algorithm,order", [(Algorithms.RK4, 4), (Algorithms.RKCK45, 5)])
def test_global_order_of_accuracy(oracle, algorithm, order):
    # a wrong tableau coefficient drops the order; halving h must divide the
    # global error by ~2^order (4 for RK4, 5 for the Cash-Karp 5th-order weights)
    t1 = 2.0
    exact = _duffing_exact(t1)
    errs = []
    for h in (0.1, 0.05, 0.025):
        o = run(oracle, _fixed_step_duffing(algorithm, h, t1))
        errs.append(np.abs(o["X"][:, 0] - exact).max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(order - 0.3 < r < order + 0.6 for r in rates), (errs, rates)


def test_embedded_error_estimate_is_fifth_order_locally(oracle):
    # err = h |sum (b - b*) k| must scale as h^5 for one step (tol = atol = 1, so ratio = 1/max err)
    ratios = []
    for h in (0.02, 0.01, 0.005):
        s = _fixed_step_duffing(Algorithms.RKCK45, h, 1.0)
        s["rtol"], s["atol"] = 0.0, 1.0
        tr = oracle.trace("duffing", s)
        ratios.append(1.0 / tr["ratio"][0] - 1e-18)
    rate = [math.log2(ratios[i] / ratios[i + 1]) for i in range(2)]
    assert all(4.8 < r < 5.3 for r in rate), rate


# ------------------------------------------------------------------ controller
def test_constant_rhs_grows_by_the_grow_limit_and_clips(oracle):
    # x0' = 1: the embedded error is ~1e-18, so every step is accepted and grows
    # by TimeStepGrowLimit (5): 0.01, 0.05, 0.25, 1.25, 6.25, then the clip to t=10
    o = run(oracle, S.affine(nt=1, b0=1.0, td=(0.0, 10.0), rtol=1e-8, atol=1e-8))
    assert o["nacc"][0] == 6 and o["nrej"][0] == 0
    assert o["T"][0] == 10.0 and o["X"][0, 0] == pytest.approx(10.0, abs=1e-13)
    # MaximumTimeStep 1: 0.01, 0.05, 0.25, nine steps of 1, the clipped 0.69
    o = run(oracle, S.affine(nt=1, b0=1.0, td=(0.0, 10.0), dtmax=1.0))
    assert o["nacc"][0] == 13


def test_controller_fixed_point(oracle):
    # x' = -x with tol = rtol*|x|: err/tol depends on h only, and the step rule
    # h <- 0.9 h (tol/err)^(1/5) has its fixed point at tol/err = 0.9^-5 exactly
    tr = oracle.trace("affine2", S.affine(nt=1, a00=-1.0, a11=-1.0, x0=(1.0, 2.0), td=(0.0, 5.0),
                                          rtol=1e-6, atol=0.0))
    acc = tr["accepted"] == 1
    r = tr["ratio"][acc][10:-1]        # converges geometrically, then sits at the fixed point
    assert len(r) >= 8 and np.abs(r / 0.9 ** -5 - 1).max() < 1e-9
    h = tr["h"][acc][10:-1]
    assert np.abs(h / h[0] - 1).max() < 1e-9


def test_rejection_uses_quarter_power_and_shrink_limit(oracle):
    # PT-CTL:51-53, 79-85: after a rejection the next attempt is
    # h * clamp(0.9 R^(1/4), shrink, grow); after an acceptance 0.9 R^(1/5)
    tr = oracle.trace("affine2", S.affine(nt=1, a00=-50.0, a11=-3.0, x0=(1.0, 1.0), td=(0.0, 1.0), dt0=1.0,
                                          rtol=1e-8, atol=1e-8))
    assert tr["accepted"][0] == 0 and tr["accepted"].sum() > 5
    for i in range(len(tr["h"]) - 2):
        e = 0.2 if tr["accepted"][i] else 0.25
        fac = min(max(0.9 * tr["ratio"][i] ** e, 0.1), 5.0)
        if tr["t"][i + 1] + tr["h"][i + 1] < 1.0 - 1e-12:    # skip the clipped last step
            assert tr["h"][i + 1] == pytest.approx(tr["h"][i] * fac, rel=1e-14)
    assert tr["h"][1] == pytest.approx(0.1)               # first retry hits the shrink limit


def test_minimum_step_forces_acceptance(oracle):
    o = run(oracle, S.affine(nt=1, a01=1.0, a10=-1.0, x0=(1.0, 0.0), td=(0.0, 0.05), dtmin=1e-3,
                             rtol=0.0, atol=0.0))
    # zero tolerance: every attempt fails the error test; the step shrinks to the
    # minimum and is then forced through (PT-CTL:71-75)
    assert o["status"][0] & 4
    assert o["nrej"][0] == 1 and o["nacc"][0] in (50, 51)
    assert o["X"][0, 0] == pytest.approx(math.cos(0.05), abs=1e-12)


def test_nonfinite_state(oracle):
    o = run(oracle, S.affine(nt=1, a00=-1.0, x0=(np.nan, 1.0), td=(0.0, 1.0), algorithm=Algorithms.RK4))
    assert o["status"][0] == 1 and o["nacc"][0] == 1 and np.isnan(o["X"][0, 0])   # committed, then terminated
    o = run(oracle, S.affine(nt=1, a00=-1.0, x0=(np.nan, 1.0), td=(0.0, 1.0)))
    assert o["status"][0] & 2 and o["nacc"][0] == 0
    # rejected with the shrink factor until dt < 1.01 dtmin: 0.01 * 0.1^k
    assert o["nrej"][0] == 11


# ------------------------------------------------------------------ events
def test_event_step_lands_on_the_crossing(oracle):
    o = run(oracle, S.event_ramp())
    assert o["nev"][0] == 1 and o["ACCi"][0, 0] == 1
    assert o["nrej"][0] == 1 and o["nacc"][0] == 5            # 0.3, 0.5 (after the retry), 0.7, 0.9, 1.0
    assert o["evlog"][0, 0, 0] == pytest.approx(0.5, abs=1e-15)
    assert o["evlog"][0, 1, 0] == pytest.approx(0.5, abs=1e-15)


def test_event_direction_and_omission(oracle):
    o = run(oracle, S.event_ramp(edir=-1))                     # increasing crossing ignored
    assert o["nev"][0] == 0 and o["nrej"][0] == 0 and o["nacc"][0] == 4
    o = run(oracle, S.event_ramp(dtmin=0.25))                  # interpolated step 0.2 < dtmin: omitted
    assert o["status"][0] & 8 and o["nev"][0] == 0 and o["nrej"][0] == 0


def test_poincare_section_times_match_scipy(oracle):
    # x' = 0 crossings of the uneven-time-scale Duffing system at eps = 1
    s = W.new_spec("duffing_eps", Algorithms.RKCK45, 1)
    s["cPAR"][:, 0] = (0.2, 2.0, 1.0)
    s["X"][:, 0] = (-0.5, 0.1)
    s["TD"][1] = 20.0
    s["rtol"] = s["atol"] = 1e-10
    s["evcap"] = 64
    o = run(oracle, s)
    n = o["nev"][0]
    f = lambda t, y: [y[1], 2.0 * math.cos(t) - 0.2 * y[1] - y[0] - y[0] ** 3]
    ev = lambda t, y: y[1]
    sol = solve_ivp(f, (0, 20), [-0.5, 0.1], method="DOP853", rtol=1e-13, atol=1e-13, events=ev)
    ref = sol.t_events[0]
    assert n == len(ref)
    # detection fires once |x'| < 1e-6: time accurate to ~1e-6 / |x''|
    assert np.abs(o["evlog"][:n, 0, 0] - ref).max() < 2e-5


# ------------------------------------------------------------------ dense output, time grid
@pytest.mark.parametrize("freq,mindt", [(1, 0.0), (3, 0.0), (1, 0.25)])
def test_dense_output_brute_force(oracle, freq, mindt):
    spec = S.fixed_step(nt=1, b0=1.0, td=(0.0, 1.0), dt0=0.1, ndo=16, do_freq=freq, do_min_dt=mindt)
    o = run(oracle, spec)
    grid = S.brute_force_grid(0.0, 1.0, 0.1)
    # PT-DO:5-58 restated on the brute-force grid
    stored, skipped, nxt = [0.0], 0, min(0.0 + mindt, 1.0)
    for i, t in enumerate(grid):
        last = i == len(grid) - 1
        if (len(stored) < 16 and nxt < t and skipped >= freq - 1) or (len(stored) < 16 and last):
            stored.append(t)
            skipped, nxt = 0, min(t + mindt, 1.0)
        else:
            skipped += 1
    n = o["DOidx"][0]
    assert n == len(stored)
    assert np.array_equal(o["DOt"][:n, 0], np.array(stored))
    assert np.allclose(o["DOx"][:n, 0, 0], np.array(stored), atol=1e-14)   # x0 = t


def test_rk4_step_counts_follow_the_clip_rule(oracle):
    # one Solve over 64 forcing periods: 64001 steps at 2pi/1000 (last one a
    # 4.8e-10 sliver), 32000 at 2pi/500; brute force over the same IEEE grid
    for h, n in ((2 * W.PI / 1000, 64001), (2 * W.PI / 500, 32000)):
        assert len(S.brute_force_grid(0.0, 64 * (2 * W.PI), h)) == n
    o = run(oracle, W.config0_duffing_rk4(nt=4))
    assert (o["nacc"] == 64001).all() and (o["T"] == 64 * (2 * W.PI)).all()


# ------------------------------------------------------------------ against SciPy, invariants
def test_lorenz_matches_dop853_outside_chaos(oracle):
    s = W.new_spec("lorenz", Algorithms.RKCK45, 4)
    rhos = np.array([0.5, 5.0, 10.0, 13.0])
    s["cPAR"][0] = rhos
    s["X"][:] = 10.0
    s["TD"][1] = 10.0
    s["rtol"] = s["atol"] = 1e-10
    o = run(oracle, s)
    for i, rho in enumerate(rhos):
        f = lambda t, y: [10 * (y[1] - y[0]), rho * y[0] - y[1] - y[0] * y[2], y[0] * y[1] - 8 / 3 * y[2]]
        ref = solve_ivp(f, (0, 10), [10.0, 10.0, 10.0], method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1]
        assert np.abs(o["X"][:, i] - ref).max() < 1e-7 * max(1.0, np.abs(ref).max())


def test_lorenz_fixed_points(oracle):
    s = W.new_spec("lorenz", Algorithms.RKCK45, 2)
    s["cPAR"][0] = (0.5, 5.0)
    s["X"][:] = 10.0
    s["TD"][1] = 200.0
    s["rtol"] = s["atol"] = 1e-10
    o = run(oracle, s)
    assert np.abs(o["X"][:, 0]).max() < 1e-8                          # rho < 1: origin
    c = math.sqrt(8 / 3 * 4.0)
    x = o["X"][:, 1]                                                    # rho = 5: C+- = (+-sqrt(b(rho-1)), ., rho-1)
    assert np.allclose(np.abs(x), [c, c, 4.0], atol=1e-8) and x[0] == pytest.approx(x[1], abs=1e-8)


def test_duffing_rk4_matches_dop853(oracle):
    s = W.config0_duffing_rk4(nt=3)
    s["TD"][1] = 2 * (2 * W.PI)
    o = run(oracle, s)
    for i in range(3):
        B = s["cPAR"][1, i]
        f = lambda t, y: [y[1], B * math.cos(t) - 0.2 * y[1] - y[0] - y[0] ** 3]
        ref = solve_ivp(f, (0, 4 * W.PI), [-0.5, 0.1], method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1]
        assert np.abs(o["X"][:, i] - ref).max() < 5e-7    # RK4 global error at h = 2pi/1000


def test_accessory_tracks_max_abs_x(oracle):
    s = W.config0_duffing_rk4(nt=2)
    s["TD"][1] = 2 * (2 * W.PI)
    o = run(oracle, s)
    B = s["cPAR"][1, 1]
    f = lambda t, y: [y[1], B * math.cos(t) - 0.2 * y[1] - y[0] - y[0] ** 3]
    sol = solve_ivp(f, (0, 4 * W.PI), [-0.5, 0.1], method="DOP853", rtol=1e-12, atol=1e-12, dense_output=True)
    grid = np.arange(0, 4 * W.PI, 2 * W.PI / 1000)
    ref = np.abs(sol.sol(grid)[0]).max()
    assert o["ACC"][0, 1] == pytest.approx(ref, abs=1e-6)


def test_tutorial_callbacks_semantics(oracle):
    # the double-well tutorial system: ACCi[0] counts x = 0 crossings, ACC[2] >= ACC[0]
    name, system, spec = [x for x in S.all_scenarios() if x[0] == "tutorial_duffing_doublewell"][0]
    o = run(oracle, spec)
    assert (o["T"] == 2 * W.PI).all()
    assert (o["DOidx"] > 10).all() and (o["DOidx"] <= 200).all()
    assert (o["ACC"][2] >= o["ACC"][0] - 1e-12).all()
    f = lambda k: (lambda t, y: [y[1], y[0] - y[0] ** 3 - k * y[1] + 0.3 * math.cos(t)])
    ref = solve_ivp(f(spec["cPAR"][0, 0]), (0, 2 * W.PI), [-0.5, -0.1], method="DOP853", rtol=1e-13, atol=1e-13,
                    events=lambda t, y: y[0])
    assert o["ACCi"][0, 0] == len(ref.t_events[0])
    assert np.abs(o["X"][:, 0] - ref.y[:, -1]).max() < 1e-7


# ------------------------------------------------------------------ forward sensitivities (adjoint reference)
def test_sensitivity_of_rk4_on_a_linear_system_is_the_matrix_power(oracle):
    # x' = A x + b with exact step grid h = 1/8 over [0, 2]: the RK4 map is
    # x <- M x + N b, M = I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24, so
    # d x(T)/d x(0) = M^16 and d x(T)/d b = sum_k M^k N
    A = np.array([[-0.5, 1.0], [-1.0, -0.1]])
    b = np.array([0.3, -0.2])
    spec = S.affine(nt=1, a00=A[0, 0], a01=A[0, 1], a10=A[1, 0], a11=A[1, 1], b0=b[0], b1=b[1], x0=(1.0, 0.5),
                    td=(0.0, 2.0), dt0=0.125, algorithm=Algorithms.RK4)
    o = oracle.sensitivity("affine2", spec)
    h = 0.125
    hA = h * A
    I = np.eye(2)
    M = I + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    N = h * (I + hA / 2 + hA @ hA / 6 + hA @ hA @ hA / 24)
    J = o["jac"][:, :, 0]
    assert np.allclose(J[:, :2], np.linalg.matrix_power(M, 16), rtol=1e-13, atol=1e-14)
    G = sum(np.linalg.matrix_power(M, k) for k in range(16)) @ N
    assert np.allclose(J[:, 6:8], G, rtol=1e-12, atol=1e-14)
    # d x(T)/d a00 by central differences of the oracle itself (smooth map: fixed step)
    eps = 1e-6
    xs = []
    for d in (eps, -eps):
        sp = S.affine(nt=1, a00=A[0, 0] + d, a01=A[0, 1], a10=A[1, 0], a11=A[1, 1], b0=b[0], b1=b[1], x0=(1.0, 0.5),
                      td=(0.0, 2.0), dt0=0.125, algorithm=Algorithms.RK4)
        xs.append(oracle.run("affine2", sp)["X"][:, 0])
    assert np.allclose((xs[0] - xs[1]) / (2 * eps), J[:, 2], rtol=1e-8, atol=1e-10)


def _relief_valve_scipy(q, kappa=1.25, delta=10.0, beta=20.0, r=0.8, x0=(0.2, 0.0, 0.0)):
    """The impact tutorial's first Solve with SciPy: integrate to the first
    downward zero of v (terminal), restarting after every downward zero of x
    with v -> -r v.  Returns (t_end, state, impacts)."""
    def f(t, y):
        return [y[1], -kappa * y[1] - (y[0] + delta) + y[2], beta * (q - y[0] * math.sqrt(max(y[2], 0.0)))]

    def impact(t, y):
        return y[0]
    impact.terminal, impact.direction = True, -1

    def peak(t, y):
        return y[1]
    peak.terminal, peak.direction = True, -1
    # v starts at 0 and falls: a crossing needs a positive value before it (as in
    # the solver's detection), so step off the initial zero first
    first = solve_ivp(f, (0.0, 1e-6), np.array(x0, float), method="DOP853", rtol=1e-13, atol=1e-13)
    t, y, impacts = 1e-6, first.y[:, -1], 0
    for _ in range(100):
        sol = solve_ivp(f, (t, t + 100.0), y, method="DOP853", rtol=1e-13, atol=1e-13, events=(impact, peak))
        if sol.t_events[1].size:
            return sol.t_events[1][0], sol.y_events[1][0], impacts
        t, y = sol.t_events[0][0], sol.y_events[0][0].copy()
        y[0] = 0.0
        y[1] = -r * y[1]
        impacts += 1
    raise AssertionError("no peak")


def test_relief_valve_impacts_and_user_termination_match_scipy(oracle):
    """Impact tutorial (state jump at x = 0, user termination at the first
    maximum of x): the oracle's end time and state agree with an event-driven
    SciPy integration to the event tolerance's effect (1e-5); instances with
    bounces before the maximum are among them."""
    spec = W.tutorial_relief_valve(nt=12)
    out = oracle.run("relief_valve", spec)
    q = spec["cPAR"][0]
    bounced = 0
    for i in range(spec["nt"]):
        t_end, y_end, impacts = _relief_valve_scipy(q[i])
        bounced += impacts > 0
        assert out["ACCi"][0, i] == 1 and out["ACCi"][1, i] < 50      # ended by the first maximum
        assert abs(out["T"][i] - t_end) < 1e-5 * max(1.0, t_end), (i, out["T"][i], t_end)
        assert np.allclose(out["X"][[0, 2], i], y_end[[0, 2]], rtol=1e-5, atol=1e-6), (i, out["X"][:, i], y_end)
        assert abs(out["X"][1, i]) < 1e-6                              # v within the event tolerance
    assert bounced >= 2


def _keller_miksis_rhs(c):
    def f(t, y):
        rx1 = 1.0 / y[0]
        p = rx1 ** c[10]
        s1, c1 = math.sin(2 * W.PI * t), math.cos(2 * W.PI * t)
        a2 = 2.0 * c[11] * W.PI * t + c[12]
        s2, c2 = math.sin(a2), math.cos(a2)
        N = ((c[0] + c[1] * y[1]) * p - c[2] * (1.0 + c[9] * y[1]) - c[3] * rx1 - c[4] * y[1] * rx1
             - 1.5 * (1.0 - c[9] * y[1] / 3.0) * y[1] ** 2 - (c[5] * s1 + c[6] * s2) * (1.0 + c[9] * y[1])
             - y[0] * (c[7] * c1 + c[8] * c2))
        D = y[0] - c[9] * y[0] * y[1] + c[4] * c[9]
        return [y[1], N / D]
    return f


def test_keller_miksis_first_maximum_matches_scipy(oracle):
    """Quasi-periodic forcing tutorial (Keller-Miksis bubble, two forcing
    frequencies, user termination at the first maximum of the radius): the
    oracle's end time and radius agree with an event-driven SciPy integration
    (DOP853 at 1e-12, terminal event y2 = 0 falling) to the event tolerance's
    effect, and the minimum-radius accessory with the SciPy minimum before it."""
    spec = W.tutorial_keller_miksis(4, 4)
    out = oracle.run("keller_miksis", spec)
    for i in range(spec["nt"]):
        c = spec["cPAR"][:, i]
        ev = lambda t, y: y[1]
        ev.direction = -1
        sol = solve_ivp(_keller_miksis_rhs(c), (0.0, 2.0 * out["T"][i] + 1.0), [1.0, 0.0], method="DOP853",
                        rtol=1e-12, atol=1e-12, events=ev, dense_output=True)
        # y2 starts at 0: the solver's event test needs y2 above the tolerance
        # first, so the start is not an event; the first maximum after it is
        k = int(np.argmax(sol.t_events[0] > 1e-9))
        t_end = sol.t_events[0][k]
        y_end = sol.y_events[0][k]
        assert out["ACCi"][0, i] == 1                                   # ended by the first maximum
        assert abs(out["T"][i] - t_end) < 1e-5 * max(1.0, t_end), (i, out["T"][i], t_end)
        assert abs(out["X"][0, i] - y_end[0]) < 1e-6 * abs(y_end[0]), (i, out["X"][0, i], y_end[0])
        assert abs(out["X"][1, i]) < 1e-5
        tt = np.linspace(0.0, t_end, 20001)
        assert out["ACC"][2, i] == pytest.approx(sol.sol(tt)[0].min(), rel=1e-4)


def test_keller_miksis_one_period_matches_scipy(oracle):
    """Large time scale differences tutorial (single forcing, one period per
    Solve): the oracle's state after one period and its largest-radius
    accessory agree with SciPy (DOP853 at 1e-12) to 1e-7 relative."""
    spec = W.tutorial_keller_miksis_max(8)
    out = oracle.run("keller_miksis_max", spec)
    for i in range(spec["nt"]):
        c = list(spec["cPAR"][:, i]) + [0.0] * 8
        sol = solve_ivp(_keller_miksis_rhs(c), (0.0, 1.0), [1.0, 0.0], method="DOP853", rtol=1e-12, atol=1e-12,
                        dense_output=True)
        assert np.allclose(out["X"][:, i], sol.y[:, -1], rtol=1e-7, atol=1e-9), (i, out["X"][:, i], sol.y[:, -1])
        ymax = sol.sol(np.linspace(0.0, 1.0, 200001))[0].max()
        assert out["ACC"][0, i] == pytest.approx(ymax, rel=1e-5)
