"""The library's elementary functions (csrc/mpgos_math.h, host build inside
libmpgos.so) against the platform libm / numpy and high-precision mpmath.
They replace the CUDA libm calls of the reference (pow in
SingleSystem_PerThread_ExplicitRungeKutta_ErrorControllers.cuh:51-53, cos in
the user right-hand sides), whose exact bits are not pinned by anything in
the reference; what is pinned here is accuracy: a few ulp everywhere the
solver uses them."""
import ctypes

import mpmath
import numpy as np
import pytest

from mpgos import lib

FN = {"cos": 0, "sin": 1, "exp": 2, "log": 3, "pow": 4, "sqrt": 5, "pow_ctrl": 6, "ctrl_root": 7}


def host_eval(name, x, y=None):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    yp = None
    if y is not None:
        y = np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float64)
        yp = y.ctypes.data
    rc = lib().mpgos_math_eval_host(FN[name], x.ctypes.data, yp, out.ctypes.data, x.size)
    assert rc == 0
    return out


def ulp_err(a, ref):
    ref = np.asarray(ref, np.float64)
    return np.abs(a - ref) / np.spacing(np.abs(ref))


def mp_ref(fn, xs, ys=None):
    mpmath.mp.dps = 40
    if ys is None:
        return np.array([float(fn(mpmath.mpf(float(v)))) for v in xs])
    return np.array([float(fn(mpmath.mpf(float(a)), mpmath.mpf(float(b)))) for a, b in zip(xs, ys)])


rng = np.random.default_rng(1)


@pytest.mark.parametrize("name,mfn", [("cos", mpmath.cos), ("sin", mpmath.sin)])
def test_trig_against_mpmath(name, mfn):
    # forcing arguments reach ~2.6e4 (uneven-time-scale workload); stress up to 1e5
    x = np.concatenate([rng.uniform(-4, 4, 400), rng.uniform(-3e4, 3e4, 400), rng.uniform(-1e5, 1e5, 200),
                        10.0 ** rng.uniform(-12, -1, 100)])
    got = host_eval(name, x)
    ref = mp_ref(mfn, x)
    big = np.abs(ref) > 1e-3
    assert ulp_err(got[big], ref[big]).max() <= 2.0
    # near zeros of the function the absolute error is what counts
    assert np.abs(got - ref).max() <= 4e-16 * max(1.0, 1e-5 * np.abs(x).max())


@pytest.mark.parametrize("name,mfn", [("cos", mpmath.cos), ("sin", mpmath.sin)])
def test_trig_large_arguments(name, mfn):
    # chained Solves and slow forcings push phases far beyond one period: every
    # decade up to 2^50 stays within a few 1e-16 of the true value (absolute)
    x = np.concatenate([10.0 ** rng.uniform(5, 15.05, 600), -(10.0 ** rng.uniform(5, 15.05, 200)),
                        [2.0 ** 50 - 0.25, -(2.0 ** 50 - 0.125), 1e15 + 0.25]])
    got = host_eval(name, x)
    with mpmath.workdps(60):
        ref = np.array([float(mfn(mpmath.mpf(float(a)))) for a in x])
    assert np.abs(got - ref).max() <= 4e-16, np.abs(got - ref).max()
    # from 2^50 on the spacing of doubles is >= 1/4 rad: NaN, so the solver
    # reports the instance as non-finite instead of integrating noise
    huge = np.array([2.0 ** 50, -(2.0 ** 50), 2.0 ** 52, 1e19, 1e300, np.finfo(np.float64).max])
    assert np.isnan(host_eval(name, huge)).all()


def test_trig_bulk_against_numpy():
    x = np.concatenate([np.linspace(-500, 500, 200001), rng.uniform(-3e4, 3e4, 200000)])
    for name, f in (("cos", np.cos), ("sin", np.sin)):
        got = host_eval(name, x)
        ref = f(x)
        assert np.abs(got - ref).max() < 5e-16 * 4, name


def test_trig_identities_and_special_values():
    x = rng.uniform(-1e3, 1e3, 10000)
    c, s = host_eval("cos", x), host_eval("sin", x)
    assert np.abs(c * c + s * s - 1).max() < 4e-16
    assert np.isnan(host_eval("cos", np.array([np.inf, -np.inf, np.nan]))).all()
    assert host_eval("cos", np.array([0.0]))[0] == 1.0
    assert host_eval("sin", np.array([-0.0, 1e-300]))[1] == 1e-300


def test_exp_log():
    x = np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-1, 1, 1000), [0.0, 1e-300, -745.0, 709.7]])
    assert ulp_err(host_eval("exp", x), mp_ref(mpmath.exp, x)).max() <= 2.0
    y = np.concatenate([10.0 ** rng.uniform(-300, 300, 2000), rng.uniform(0.5, 2.0, 1000),
                        [5e-324, 2.2250738585072014e-308, 1.0 + 2**-52, 1e308]])
    ly = host_eval("log", y)
    ref = mp_ref(mpmath.log, y)
    ok = np.abs(ref) > 1e-3
    assert ulp_err(ly[ok], ref[ok]).max() <= 2.0
    assert np.abs(ly - ref).max() < 1e-15 * np.abs(ref).max()
    assert host_eval("log", np.array([1.0]))[0] == 0.0
    assert np.isnan(host_eval("log", np.array([-1.0, np.nan]))).all()
    assert host_eval("log", np.array([0.0]))[0] == -np.inf
    assert host_eval("exp", np.array([1000.0]))[0] == np.inf
    assert host_eval("exp", np.array([-1000.0]))[0] == 0.0


def test_pow_controller_range():
    # the controller evaluates R^(1/5) and R^(1/4) for R = min(tol/err) in [0, 1e30]
    r = np.concatenate([10.0 ** rng.uniform(-20, 30, 3000), [1.0, 1e30, 0.9 ** -5]])
    for e in (0.2, 0.25):
        got = host_eval("pow", r, e)
        ref = mp_ref(mpmath.power, r, np.full_like(r, e))
        assert ulp_err(got, ref).max() <= 4.0, e
    assert host_eval("pow", np.array([0.0]), 0.2)[0] == 0.0
    assert host_eval("pow", np.array([np.inf]), 0.2)[0] == np.inf
    assert np.isnan(host_eval("pow", np.array([np.nan]), 0.2)[0])
    assert host_eval("pow", np.array([-2.0]), 3.0)[0] == -8.0


def test_float_versions_round_once():
    # FP32 mode: float functions are the double ones rounded once
    x = rng.uniform(-100, 100, 1000).astype(np.float32).astype(np.float64)
    assert np.array_equal(host_eval("cos", x).astype(np.float32), np.cos(x).astype(np.float32)) or \
        np.abs(host_eval("cos", x).astype(np.float32) - np.cos(x).astype(np.float32)).max() <= 6e-8


def test_sqrt_is_correctly_rounded():
    rng = np.random.default_rng(11)
    x = np.concatenate([10.0 ** rng.uniform(-308, 308, 20000), rng.uniform(0, 10, 20000),
                        np.array([0.0, -0.0, np.inf, 5e-324, 2.2250738585072014e-308, 1.0, 4.0])])
    out = host_eval("sqrt", x)
    assert np.array_equal(out.view(np.uint64), np.sqrt(x).view(np.uint64))
    assert np.isnan(host_eval("sqrt", np.array([-1.0, np.nan]))).all()


def _pow_rules(x, y):
    """pow's special cases, in order of precedence, then the general value
    (mpmath on |x|, signed by the exponent's parity for a negative base)."""
    import math
    if math.isnan(x) or math.isnan(y):
        return math.nan
    if y == 0.0 or x == 1.0:
        return 1.0
    if x == 0.0:
        return 0.0 if y > 0.0 else math.inf
    if math.isinf(x):
        return (x if y > 0.0 else 0.0) if x > 0.0 else math.nan
    sign = 1.0
    if x < 0.0:
        if float(np.rint(y)) != y:
            return math.nan
        if float(np.rint(0.5 * y)) != 0.5 * y:
            sign = -1.0
    if math.isinf(y):   # |x| = 1: inf * log 1 is NaN (C99 would give 1)
        if abs(x) == 1.0:
            return math.nan
        return sign * (math.inf if (abs(x) > 1.0) == (y > 0) else 0.0)
    mpmath.mp.dps = 40
    v = mpmath.power(mpmath.mpf(abs(x)), mpmath.mpf(y))
    return sign * (math.inf if v > 1.8e308 else float(v))


def test_pow_and_exp_special_cases():
    """Every combination of special and ordinary arguments against the rules
    above: zeros of both signs, infinities, NaN, one, negative bases with
    integer, odd and fractional exponents, subnormals, over- and underflow."""
    xs = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, 1e-310, -3.0, 1e30, 7.25]
    ys = [0.0, -0.0, 1.0, 2.0, 3.0, -3.0, 0.5, -0.5, 0.2, 0.25, np.inf, -np.inf, np.nan, 1e300, -1e300, 2.5]
    X, Y = np.meshgrid(np.array(xs), np.array(ys))
    got = host_eval("pow", X.ravel(), Y.ravel())
    for x, y, g in zip(X.ravel(), Y.ravel(), got):
        want = _pow_rules(float(x), float(y))
        if np.isnan(want):
            assert np.isnan(g), (x, y, g)
        elif want == 0.0 or np.isinf(want):
            assert g == want, (x, y, g, want)
        else:   # ordinary values: accuracy is pinned in the controller's range above
            assert abs(g - want) <= max(1e-14 * abs(want), 4 * 5e-324), (x, y, g, want)
    hi, lo = 709.782712893384, -745.1332191019412
    xe = np.array([np.nan, np.inf, -np.inf, hi, np.nextafter(hi, np.inf), lo, np.nextafter(lo, -np.inf),
                   -744.0, -708.5, 0.0, -0.0])
    ge = host_eval("exp", xe)
    assert np.isnan(ge[0]) and ge[1] == np.inf and ge[2] == 0.0
    assert np.isfinite(ge[3]) and ge[4] == np.inf and ge[5] in (0.0, 5e-324) and ge[6] == 0.0
    ref = mp_ref(mpmath.exp, xe[[3, 7, 8, 9, 10]])
    assert ulp_err(ge[[3, 7, 8, 9, 10]], ref).max() <= 2.0


def test_controller_power_equals_pow_on_its_domain():
    """pow_ctrl (the controllers' R^(1/5), R^(1/4)) returns pow's bits for
    x >= 0, +inf or NaN and 0 < y <= 1/4: random x over the whole positive range
    including subnormals, the special arguments, and random exponents."""
    x = np.concatenate([10.0 ** rng.uniform(-323, 308, 200000), rng.uniform(0.0, 2.0, 100000),
                        [0.0, -0.0, np.inf, np.nan, 1.0, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                         0.9 ** -5, 1e30]])
    for y in (np.full(x.size, 0.2), np.full(x.size, 0.25), rng.uniform(1e-3, 0.25, x.size)):
        a = host_eval("pow_ctrl", x, y)
        b = host_eval("pow", x, y)
        nan = np.isnan(a) & np.isnan(b)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a.view(np.uint64)[~nan], b.view(np.uint64)[~nan])


def test_controller_root_against_mpmath():
    """ctrl_root (the per-thread controller's R^(1/5) after an accepted step,
    R^(1/4) after a rejection, from a start polynomial, two Newton steps and
    square roots) against mpmath over the whole non-negative range, subnormals
    and the controller's fixed point 0.9^-5 included: within 2 ulp for the
    fifth root, 1 ulp for the fourth; 0, +inf and NaN pass through."""
    x = np.concatenate([10.0 ** rng.uniform(-323, 308, 4000), rng.uniform(0.0, 2.0, 4000), rng.uniform(0.9, 1.1, 2000),
                        [5e-324, 2.2250738585072014e-308, 2.0 ** -1000, 2.0 ** -1001, 1.0, 32.0, 0.9 ** -5, 1e30,
                         1.7976931348623157e308]])
    for fifth, e in ((1.0, mpmath.mpf(1) / 5), (0.0, mpmath.mpf(1) / 4)):
        got = host_eval("ctrl_root", x, fifth)
        ref = mp_ref(lambda v: v ** e, x)
        assert ulp_err(got, ref).max() <= (2.0 if fifth else 1.0), (fifth, ulp_err(got, ref).max())
        special = host_eval("ctrl_root", np.array([0.0, np.inf, np.nan]), fifth)
        assert special[0] == 0.0 and special[1] == np.inf and np.isnan(special[2])
    # the fixed point of h <- 0.9 h (tol/err)^(1/5) stays a fixed point
    assert abs(0.9 * host_eval("ctrl_root", np.array([0.9 ** -5]), 1.0)[0] - 1.0) <= 2.3e-16
