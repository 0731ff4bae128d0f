"""Small integration scenarios whose outcome mathematics fixes (closed forms,
invariants, brute-force time grids).  The oracle is pinned against them in
test_oracle_pins.py; the GPU runs the same specs in test_gpu_parity.py and
must match the oracle bit for bit."""
import numpy as np

from mpgos import Algorithms
from mpgos import workloads as W


def affine(nt, a00=0.0, a01=0.0, a10=0.0, a11=0.0, b0=0.0, b1=0.0, x0=(0.0, 0.0), td=(0.0, 1.0),
           algorithm=Algorithms.RKCK45, ndo=0, c=1e30, **options):
    """affine2: x' = A x + b, event EF = x0 - c (detections counted in ACCi[0])."""
    s = W.new_spec("affine2", algorithm, nt, ndo=ndo)
    for i, v in enumerate((a00, a01, a10, a11, b0, b1)):
        s["cPAR"][i] = v
    s["X"][0], s["X"][1] = x0
    s["TD"][0], s["TD"][1] = td
    s["T"][:] = td[0]
    s["sPAR"][0] = c
    for k in ("rtol", "atol", "etol", "edir", "evcap"):
        if k in options:
            s[k] = options.pop(k)
    s["options"].update(options)
    return s


def fixed_step(**kw):
    """RKCK45 with grow = shrink = 1: every accepted step keeps the initial size."""
    kw.setdefault("grow", 1.0)
    kw.setdefault("shrink", 1.0)
    kw.setdefault("rtol", 1e3)
    kw.setdefault("atol", 1e3)
    return affine(**kw)


def brute_force_grid(t0, t1, h):
    """Accepted-step times of a fixed step h with the reference's clip rule
    (clip when h > t1 - t, stop after the clipped step; PT-SOL:204-210, 391-392)."""
    t, times = t0, []
    while True:
        dt = h
        end = False
        if dt > t1 - t:
            dt = t1 - t
            end = True
        t = t + dt
        times.append(t)
        if end:
            return np.array(times)


def event_ramp(nt=1, edir=0, dtmin=1e-12):
    """x0' = 1 from 0 with fixed steps 0.3 and the event x0 = 0.5: the step
    into (0.5, 0.6] is rejected, shortened by linear interpolation to exactly
    the crossing, then the event is detected at t = 0.5."""
    return fixed_step(nt=nt, b0=1.0, x0=(0.0, 0.0), td=(0.0, 1.0), c=0.5, dt0=0.3, edir=edir, etol=1e-6,
                      dtmin=dtmin, evcap=4)


def all_scenarios():
    """(name, system, spec) of every scenario, for the GPU parity sweep."""
    out = []
    out.append(("rk4_exp_single_step", "affine2",
                affine(nt=64, a00=-1.3, a11=0.7, x0=(1.0, -2.0), td=(0.0, 0.1), dt0=0.2,
                       algorithm=Algorithms.RK4)))
    out.append(("rkck45_constant_grow", "affine2", affine(nt=64, b0=1.0, td=(0.0, 10.0), rtol=1e-8, atol=1e-8)))
    out.append(("rkck45_constant_dtmax", "affine2", affine(nt=64, b0=1.0, td=(0.0, 10.0), dtmax=1.0)))
    out.append(("rkck45_decay_fixed_point", "affine2",
                affine(nt=64, a00=-1.0, a11=-1.0, x0=(1.0, 2.0), td=(0.0, 5.0), rtol=1e-6, atol=0.0)))
    out.append(("rkck45_stiff_rejections", "affine2",
                affine(nt=64, a00=-50.0, a11=-3.0, x0=(1.0, 1.0), td=(0.0, 1.0), dt0=1.0, rtol=1e-8, atol=1e-8)))
    out.append(("rkck45_min_step", "affine2",
                affine(nt=64, a01=1.0, a10=-1.0, x0=(1.0, 0.0), td=(0.0, 0.05), dtmin=1e-3, rtol=0.0, atol=0.0)))
    out.append(("event_ramp_both", "affine2", event_ramp(nt=64, edir=0)))
    out.append(("event_ramp_down_only", "affine2", event_ramp(nt=64, edir=-1)))
    out.append(("event_ramp_omitted", "affine2", event_ramp(nt=64, edir=0, dtmin=0.25)))
    nan = affine(nt=64, a00=-1.0, x0=(np.nan, 1.0), td=(0.0, 1.0))
    out.append(("rkck45_nonfinite", "affine2", nan))
    nan4 = affine(nt=64, a00=-1.0, x0=(np.nan, 1.0), td=(0.0, 1.0), algorithm=Algorithms.RK4)
    out.append(("rk4_nonfinite", "affine2", nan4))
    for freq, mindt in ((1, 0.0), (3, 0.0), (1, 0.25)):
        out.append((f"dense_output_f{freq}_m{mindt}", "affine2",
                    fixed_step(nt=64, b0=1.0, td=(0.0, 1.0), dt0=0.1, ndo=16, do_freq=freq, do_min_dt=mindt)))
    # reference tutorial system: two events, accessories, shared parameter, dense output
    dw = W.new_spec("duffing_dw", Algorithms.RKCK45, 256, ndo=200)
    dw["TD"][1] = 2 * W.PI
    dw["X"][0], dw["X"][1] = -0.5, -0.1
    dw["cPAR"][0] = W.linspace(0.2, 0.3, 256)
    dw["sPAR"][0] = 0.3
    dw["rtol"] = dw["atol"] = 1e-9
    dw["edir"] = np.array([-1, 0])
    dw["options"].update(dtmax=1e3, dtmin=1e-14, grow=10.0, shrink=0.2)
    out.append(("tutorial_duffing_doublewell", "duffing_dw", dw))
    # reference impact tutorial: state jump at impact, user termination (first maximum, or resting valve)
    out.append(("tutorial_relief_valve", "relief_valve", W.tutorial_relief_valve(nt=256)))
    # reference quasi-periodic forcing tutorial: Keller-Miksis bubble, pow and two
    # forcing frequencies in the right-hand side, termination at the first maximum
    out.append(("tutorial_keller_miksis", "keller_miksis", W.tutorial_keller_miksis(16, 16)))
    # reference large time scale differences tutorial: one period per Solve, time reset, max accessory
    out.append(("tutorial_keller_miksis_max", "keller_miksis_max", W.tutorial_keller_miksis_max(64)))
    # ragged: only part of the instances active
    lz = W.config1_lorenz(nt=1000)
    lz["n_active"] = 777
    out.append(("lorenz_ragged_active", "lorenz", lz))
    return out
