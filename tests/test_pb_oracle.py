"""Pins of the per-block (coupled systems) oracle against what the mathematics
fixes: closed-form fixed-step RK4 solutions of linear rings for every
coupling-matrix storage form, the controller's fixed point with the per-block
safety factor 0.8, SciPy's DOP853 on a dense nonlinear ring, and the event
termination time of the reference tutorial's double-well ring.

Reference semantics pinned here: coupling terms
(CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh:2634-2713), the
per-block RK4 final stage, the RKCK45 controller
(CoupledSystems_PerBlock_ExplicitRungeKutta_ErrorControllers.cuh:424-542)
and event handling (CoupledSystems_PerBlock_EventHandling.cuh:215-268).
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.optimize import brentq

from mpgos import workloads as W
from mpgos.enums import Algorithms

FORMS = [(0, 0), (0, 2), (1, 2), (1, 0)]      # (cci, cbw): full, banded, circular banded, full circular


def _matrix_for(form, ups, rs):
    """A random UPS x UPS coupling matrix that the storage form can hold
    exactly: zero outside the (wrapped) band, constant along wrapped
    diagonals for the circular forms; asymmetric, so a transpose shows."""
    cci, cbw = form
    W_ = 2 * cbw + 1 if cbw > 0 else ups
    if cci == 0:
        M = rs.normal(ups * ups).reshape(ups, ups) * 0.3
        if cbw > 0:
            for r in range(ups):
                for c in range(ups):
                    if (c - r + cbw) % ups >= W_:
                        M[r, c] = 0.0
        return M
    diag = rs.normal(W_) * 0.3
    M = np.zeros((ups, ups))
    for r in range(ups):
        for d in range(W_):
            M[r, (r + d - cbw) % ups] = diag[d]
    return M


def _linear_spec(form, ns=3, ups=12, n=64, h=1.0 / 64.0, seed=7, ndo=0):
    cci, cbw = form
    rs = W.Stream(W.SEED, 100 + seed)
    s = W.new_pb_spec("pb_linear", Algorithms.RK4, ns, ups, cbw=cbw, cci=cci, ndo=ndo)
    M = _matrix_for(form, ups, rs)
    s["cMat"] = W.store_coupling(M, cbw, cci)
    s["uPAR"][0] = rs.uniform(ns * ups, -1.0, 0.5)
    s["X"][0] = rs.uniform(ns * ups, -1.0, 1.0)
    s["strength"][0] = rs.uniform(ns, 0.5, 1.5)
    s["TD"][1] = n * h
    s["options"]["dt0"] = h
    return s, M


def _rk4_poly(Z):
    I = np.eye(Z.shape[0])
    Z2 = Z @ Z
    return I + Z + Z2 / 2 + Z2 @ Z / 6 + Z2 @ Z2 / 24


@pytest.mark.parametrize("form", FORMS, ids=["full", "banded", "circular_banded", "full_circular"])
def test_store_coupling_round_trip(form):
    """The stored form applied as the kernel indexes it equals M @ x."""
    ups = 10
    rs = W.Stream(W.SEED, 90)
    M = _matrix_for(form, ups, rs)
    cci, cbw = form
    stored = W.store_coupling(M, cbw, cci)
    assert stored.size == W.coupling_storage_size(ups, cbw, cci)
    x = rs.normal(ups)
    W_ = 2 * cbw + 1 if cbw > 0 else ups
    for u in range(ups):
        if cci == 0 and cbw == 0:
            v = sum(stored[u + c * ups] * x[c] for c in range(ups))
        else:
            v = sum((stored[u + d * ups] if cci == 0 else stored[d]) * x[(u + d - cbw) % ups] for d in range(W_))
        assert v == pytest.approx(M[u] @ x, rel=1e-13, abs=1e-14)


@pytest.mark.parametrize("form", FORMS, ids=["full", "banded", "circular_banded", "full_circular"])
def test_linear_ring_rk4_closed_form(oracle, form):
    """x' = (diag(a) + s M) x with fixed-step RK4: x_n = R(hA)^n x_0 with
    R(Z) = I + Z + Z^2/2 + Z^3/6 + Z^4/24, for every storage form; the dense
    output holds x_k at t = k h.  The end test is strict (clip when
    dt > TD[1] - t), so a step that lands exactly on TD[1] is followed by one
    zero-length clipped step, which is accepted and stored."""
    import pyoracle
    n, h = 64, 1.0 / 64.0
    s, M = _linear_spec(form, n=n, h=h, ndo=n + 2)
    o = pyoracle.run_pb("pb_linear", s)
    ups = s["ups"]
    for sid in range(s["ns"]):
        sl = slice(sid * ups, (sid + 1) * ups)
        A = np.diag(s["uPAR"][0, sl]) + s["strength"][0, sid] * M
        R = _rk4_poly(h * A)
        x = s["X"][0, sl].copy()
        for k in range(n + 1):
            assert np.allclose(o["DOx"][k, 0, sl], x, rtol=1e-12, atol=1e-14), (sid, k)
            assert o["DOt"][k, sid] == pytest.approx(k * h, abs=1e-15)
            x = R @ x
        x_end = np.linalg.matrix_power(R, n) @ s["X"][0, sl]
        assert np.array_equal(o["DOx"][n + 1, 0, sl], o["DOx"][n, 0, sl]) and o["DOt"][n + 1, sid] == n * h
        assert np.allclose(o["X"][0, sl], x_end, rtol=1e-12, atol=1e-14)
        # the same polynomial with the transposed matrix is far away: the pin sees orientation
        Rt = _rk4_poly(h * (np.diag(s["uPAR"][0, sl]) + s["strength"][0, sid] * M.T))
        if form != (1, 0) or not np.allclose(M, M.T):
            assert np.max(np.abs(np.linalg.matrix_power(Rt, n) @ s["X"][0, sl] - x_end)) > 1e-6
    assert (o["nacc"] == n + 1).all() and (o["nrej"] == 0).all()
    assert (o["DOidx"] == n + 2).all()
    assert np.array_equal(o["T"], np.full(s["ns"], n * h))


def _cash_karp_error_poly(z):
    """e(z) with h * sum (b_i - b*_i) k_i = x e(z) for x' = a x, z = h a."""
    A = np.zeros((6, 6))
    A[1, 0] = 1 / 5
    A[2, :2] = (3 / 40, 9 / 40)
    A[3, :3] = (3 / 10, -9 / 10, 6 / 5)
    A[4, :4] = (-11 / 54, 5 / 2, -70 / 27, 35 / 27)
    A[5, :5] = (1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096)
    b = np.array([37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771])
    bs = np.array([2825 / 27648, 0, 18575 / 48384, 13525 / 55296, 277 / 14336, 1 / 4])
    P = np.zeros(6)
    for i in range(6):
        P[i] = 1.0 + z * (A[i, :i] @ P[:i])
    return z * ((b - bs) @ P)


def test_rkck45_controller_fixed_point_safety_08(oracle):
    """x' = -x (one unit, no coupling), rtol 1e-8, atol 0: after the first
    steps the controller sits where 0.8 (rtol/|e(z)|)^(1/5) = 1, so the number
    of accepted steps over [0, L] is L/h* within a few; a safety factor of
    0.9 (the per-thread value) would take about 11% fewer steps."""
    import pyoracle
    s = W.new_pb_spec("pb_linear", Algorithms.RKCK45, 1, 1)
    s["uPAR"][0] = -1.0
    s["X"][0] = 1.0
    L = 20.0
    s["TD"][1] = L
    s["rtol"], s["atol"] = 1e-8, 0.0
    o = pyoracle.run_pb("pb_linear", s)

    def h_star(safety):
        target = safety ** 5 * 1e-8
        lo, hi = 1e-6, 1.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if abs(_cash_karp_error_poly(-mid)) < target:
                lo = mid
            else:
                hi = mid
        return lo

    n08 = L / h_star(0.8)
    n09 = L / h_star(0.9)
    assert n09 < 0.92 * n08
    assert abs(int(o["nacc"][0]) - n08) <= 4, (int(o["nacc"][0]), n08)
    assert o["X"][0, 0] == pytest.approx(np.exp(-L), rel=1e-6)


def test_ring_duffing_dense_coupling_vs_dop853(oracle):
    """A 16-unit forced Duffing ring with a dense asymmetric coupling matrix on
    the velocity equation, RKCK45 at 1e-11, against SciPy DOP853 at 1e-13."""
    import pyoracle
    ns, ups = 2, 16
    s = W.config4_ring(ns=ns, ups=ups, periods=1)
    s["rtol"] = s["atol"] = 1e-11
    o = pyoracle.run_pb("ring_duffing", s)
    C = s["dense_coupling"]
    for sid in range(ns):
        sl = slice(sid * ups, (sid + 1) * ups)
        k, B, cs = s["uPAR"][0, sl], s["sPAR"][0, sid], s["strength"][0, sid]

        def f(t, y):
            x, v = y[:ups], y[ups:]
            return np.concatenate([v, -k * v + B * np.cos(t) - x - x ** 3 + cs * (C @ x)])

        y0 = np.concatenate([s["X"][0, sl], s["X"][1, sl]])
        ref = solve_ivp(f, (0.0, s["TD"][1, sid]), y0, method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1]
        got = np.concatenate([o["X"][0, sl], o["X"][1, sl]])
        assert np.max(np.abs(got - ref)) < 2e-8
        # the transposed coupling gives a visibly different trajectory
        def ft(t, y):
            x, v = y[:ups], y[ups:]
            return np.concatenate([v, -k * v + B * np.cos(t) - x - x ** 3 + cs * (C.T @ x)])
        wrong = solve_ivp(ft, (0.0, s["TD"][1, sid]), y0, method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1]
        assert np.max(np.abs(wrong - ref)) > 1e-5
    assert (o["T"] == s["TD"][1]).all()


def test_tutorial_ring_dw_event_termination(oracle):
    """The reference tutorial's ring: each system stops at the second local
    maximum (x' = 0 crossing downwards) of any unit; SciPy with the same
    terminal rule fixes the stop time and state; system accessories hold
    the step count and the extreme step sizes."""
    import pyoracle
    s = W.tutorial_ring_dw()
    sids = [0, 7, 12, 24]
    sub = W.new_pb_spec("ring_dw", Algorithms.RKCK45, len(sids), s["ups"], cbw=1, cci=1, ndo=s["ndo"])
    ups = s["ups"]
    for k in ("rtol", "atol", "edir", "options", "gPAR", "cIndex", "cMat"):
        sub[k] = s[k]
    for j, sid in enumerate(sids):
        for key in ("TD", "T", "sPAR", "sACC", "sACCi", "strength", "DOidx"):
            sub[key][..., j] = s[key][..., sid]
        for key in ("X", "uPAR", "uACC", "uACCi"):
            sub[key][:, j * ups:(j + 1) * ups] = s[key][:, sid * ups:(sid + 1) * ups]
    o = pyoracle.run_pb("ring_dw", sub)
    g = sub["gPAR"][0]
    for j, sid in enumerate(sids):
        sl = slice(j * ups, (j + 1) * ups)
        om, kk, BB = (sub["uPAR"][i, sl] for i in range(3))
        cs = sub["strength"][:, j]

        def f(t, y):
            x, v = y[:ups], y[ups:]
            cx = np.roll(x, 1) + np.roll(x, -1)
            cv = np.roll(v, 1) + np.roll(v, -1)
            return np.concatenate([v - g * x + cs[0] * cx,
                                   -kk * v + x - x ** 3 + BB * np.cos(om * t) - g * v + cs[1] * cv])

        y0 = np.concatenate([sub["X"][0, sl], sub["X"][1, sl]])
        r = solve_ivp(f, (0.0, sub["TD"][1, j]), y0, method="DOP853", rtol=1e-12, atol=1e-12, dense_output=True)
        grid = np.linspace(0.0, sub["TD"][1, j], 200001)
        V = r.sol(grid)[ups:]
        second = []                       # time of each unit's second downward crossing of x' = 0
        for u in range(ups):
            idx = np.where((V[u, :-1] > 0) & (V[u, 1:] <= 0))[0]
            if len(idx) >= 2:
                i = idx[1]
                second.append((brentq(lambda t: r.sol(t)[ups + u], grid[i], grid[i + 1], xtol=1e-14), u))
        stop_t = min(second)[0]
        stop_units = [u for t, u in second if t < stop_t + 1e-9]
        stop_unit = min(stop_units, key=lambda u: abs(o["X"][1, j * ups + u]))
        assert o["T"][j] == pytest.approx(stop_t, abs=1e-6)
        assert o["uACCi"][0, j * ups + stop_unit] == 2
        assert o["uACC"][2, j * ups + stop_unit] == o["X"][0, j * ups + stop_unit]
        assert abs(o["X"][1, j * ups + stop_unit]) < sub["etol"]
        assert o["sACCi"][0, j] == o["nacc"][j]
        assert 0 < o["sACC"][0, j] <= o["sACC"][1, j]
        # dense output: the initial point plus accepted steps at least 1e-6
        # apart; the terminating step is always stored
        n_do = o["DOidx"][j]
        assert 0.9 * o["nacc"][j] < n_do <= min(sub["ndo"], o["nacc"][j] + 1)
        dts = np.diff(o["DOt"][:n_do, j])
        assert (dts[:-1] > 1e-6).all() and dts[-1] > 0
        assert o["DOt"][n_do - 1, j] == o["T"][j]
        assert np.array_equal(o["DOx"][n_do - 1, :, sl], o["X"][:, sl])


def test_host_reference_values(oracle):
    """End states that the reference solver itself produced (its device code
    compiled for the host): the per-thread T1 double-well Duffing at k = 0.2
    and the per-block T1 ring with 8 identical units.  Agreement to 1e-13
    (the reference build used glibc cos/pow, this oracle its own routines)."""
    import json
    import os
    import pyoracle
    import scenarios
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "host_reference_values.json")))
    pt = gold["per_thread_T1_duffing_doublewell"]
    spec = dict(next(s for n, _, s in scenarios.all_scenarios() if n == "tutorial_duffing_doublewell"))
    spec = W.subset(spec, [0])
    assert spec["cPAR"][0, 0] == 0.2
    o = pyoracle.run("duffing_dw", spec)
    assert abs(o["X"][0, 0] - pt["x_end"]) < 1e-13
    pb = gold["per_block_T1_ring_8_units"]
    s = W.tutorial_ring_dw(ns=1, ups=8)
    s["rtol"] = s["atol"] = 1e-10
    o = pyoracle.run_pb("ring_dw", s)
    assert o["T"][0] == pytest.approx(pb["t_end"], abs=1e-5)
    assert np.abs(o["X"][0] - pb["x0_end"]).max() < 1e-13


def test_pb_sensitivity_linear_closed_form(oracle):
    """Frozen-step sensitivities of the linear ring under fixed-step RK4:
    d x_N / d x_0(unit 0) is column 0 of R(hA)^N, and d x_N / d a_0 matches
    the complex-step derivative of the closed form."""
    import pyoracle
    n, h = 32, 1.0 / 32.0
    s, M = _linear_spec((0, 0), ns=2, ups=6, n=n, h=h)
    jac = pyoracle.sensitivity_pb("pb_linear", s)           # [ns, N=2, UD=1, U]
    ups = s["ups"]
    for sid in range(2):
        sl = slice(sid * ups, (sid + 1) * ups)
        a = s["uPAR"][0, sl].copy()

        def end_state(a_vec):
            A = np.diag(a_vec) + s["strength"][0, sid] * M
            R = _rk4_poly(h * A)
            return np.linalg.matrix_power(R, n + 1 - 1) @ s["X"][0, sl], np.linalg.matrix_power(R, n)

        _, Rn = end_state(a)
        assert np.allclose(jac[sid, 0, 0], Rn[:, 0], rtol=1e-11, atol=1e-13)
        # complex step: d/da_0 of the closed form to machine precision
        ac = a.astype(complex)
        ac[0] += 1e-30j
        cs = end_state(ac)[0].imag / 1e-30
        assert np.allclose(jac[sid, 1, 0], cs, rtol=1e-10, atol=1e-13)


def test_pb_sensitivity_ring_vs_finite_differences(oracle):
    """The dense Duffing ring with a quadrature state under RK4 (fixed steps,
    so frozen and free step sequences agree): sensitivities with respect to
    B, the initial state of unit 0 and k of unit 0 against central differences
    (h = 1e-6) of the oracle itself."""
    import pyoracle
    s = W.config4_ring(ns=2, ups=16, periods=1, algorithm=Algorithms.RK4, system="ring_duffing_q")
    s["options"]["dt0"] = 2.0 * W.PI / 200.0
    jac = pyoracle.sensitivity_pb("ring_duffing_q", s)      # [2, 5, 3, 16]
    e = 1e-6

    def fd(mod):
        sp, sm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}, \
                 {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}
        mod(sp, +e)
        mod(sm, -e)
        xp = pyoracle.run_pb("ring_duffing_q", sp)["X"]
        xm = pyoracle.run_pb("ring_duffing_q", sm)["X"]
        return (xp - xm) / (2 * e)                          # [3, ns*U]

    def dB(t, d):
        t["sPAR"][0, :] += d
    def dx0(c):
        def f(t, d):
            t["X"][c, [0, 16]] += d
        return f
    def dk(t, d):
        t["uPAR"][0, [0, 16]] += d
    cases = [(0, dB)] + [(1 + c, dx0(c)) for c in range(3)] + [(4, dk)]
    for i, mod in cases:
        ref = fd(mod)
        for sid in range(2):
            got = jac[sid, i]
            want = ref[:, sid * 16:(sid + 1) * 16]
            scale = max(1.0, np.abs(want).max())
            assert np.abs(got - want).max() < 2e-7 * scale, (i, sid)
