"""Discrete adjoint of the per-block integration on the GPU against the
oracle's frozen-step forward sensitivities (dual numbers, pinned on the CPU by
closed forms and finite differences), central finite differences of the GPU
forward for fixed-step RK4, checkpoint-interval independence, overflow
reporting, every coupling storage form (transposed coupling product), and a
two-segment mean-of-x^2 loss through a torch.autograd.Function."""
import numpy as np
import pytest

from mpgos import workloads as W
from mpgos.enums import Algorithms
from mpgos.perblock import PB, PBOption, PBProblemSolver

pytestmark = pytest.mark.gpu


def _solver(spec, kernel=None):
    s = PBProblemSolver(spec["system"], spec["ns"], spec["ups"], spec["algorithm"], spb=spec.get("spb", 1),
                        cbw=spec["cbw"], cci=spec["cci"], ndo=spec["ndo"])
    W.load_pb(s, spec)
    if kernel is not None:
        s.SolverOption(PBOption.CouplingKernel, kernel)
    return s


def _steps(spec):
    import pyoracle
    return int(pyoracle.run_pb(spec["system"], spec)["nacc"].max())


def _adjoint(spec, w, K, kernel=None):
    n = _steps(spec)
    s = _solver(spec, kernel)
    try:
        lam0, gU, gS = s.adjoint(w, K, n // K + 2, n + 2)
        s.SynchroniseFromDeviceToHost(PB.All)      # the recorded forward result
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        out = W.read_pb(s)
    finally:
        s.close()
    return lam0, gU, gS, out


def _expected(spec, jac, w):
    """Loss L = sum w * x(end): gradients from the sensitivities [ns, N, UD, U]."""
    ns, ups, ud, nsp, nup = spec["ns"], spec["ups"], spec["ud"], spec["nsp"], spec["nup"]
    wr = w.reshape(ud, ns, ups).transpose(1, 0, 2)                    # [ns, UD, U]
    g = np.einsum("sncu,scu->sn", jac, wr)                             # [ns, N]
    return g[:, :nsp], g[:, nsp:nsp + ud], g[:, nsp + ud:]


def _ring(alg, ns=10, ups=32, periods=1):
    s = W.config4_ring(ns=ns, ups=ups, periods=periods, algorithm=alg, system="ring_duffing_q")
    if alg == Algorithms.RK4:
        s["options"]["dt0"] = 2.0 * W.PI / 300.0
    return s


@pytest.mark.parametrize("kernel", [1, 2], ids=["lane_record", "mfma_record"])
@pytest.mark.parametrize("alg", [Algorithms.RK4, Algorithms.RKCK45], ids=["rk4", "rkck45"])
def test_ring_adjoint_matches_forward_sensitivities(oracle, alg, kernel):
    import pyoracle
    spec = _ring(alg)
    rs = W.Stream(W.SEED, 300)
    w = rs.normal(3 * spec["ns"] * spec["ups"]).reshape(3, -1)
    print(f"ring adjoint alg={int(alg)} kernel={kernel}", flush=True)
    lam0, gU, gS, out = _adjoint(spec, w, K=16, kernel=kernel)
    ref = pyoracle.run_pb("ring_duffing_q", spec)
    assert np.array_equal(out["X"], ref["X"]) and np.array_equal(out["nacc"], ref["nacc"])
    jac = pyoracle.sensitivity_pb("ring_duffing_q", spec)
    gB, gx0, gk0 = _expected(spec, jac, w)
    ups = spec["ups"]
    assert np.allclose(gS[0], gB[:, 0], rtol=1e-9, atol=1e-12)
    for c in range(3):
        assert np.allclose(lam0[c, ::ups], gx0[:, c], rtol=1e-9, atol=1e-12)
    assert np.allclose(gU[0, ::ups], gk0[:, 0], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("alg", [Algorithms.RK4, Algorithms.RKCK45], ids=["rk4", "rkck45"])
def test_ring_adjoint_mfma_reverse_equals_lane_reverse(oracle, alg):
    """The dense-contraction reverse pass (MFMA forward and transposed coupling
    products, stage buffer) returns the lane-per-pair reverse pass's gradients
    bit for bit; 13 systems (a partial last workgroup of 8), UPS = 64, a
    checkpoint interval that does not divide the step counts."""
    spec = _ring(alg, ns=13, ups=64)
    rs = W.Stream(W.SEED, 304)
    spec["sPAR"][0] = rs.uniform(spec["ns"], 0.5, 5.0)
    w = rs.normal(3 * spec["ns"] * spec["ups"]).reshape(3, -1)
    print(f"ring adjoint lane vs mfma alg={int(alg)}", flush=True)
    lane = _adjoint(spec, w, K=7, kernel=1)
    mfma = _adjoint(spec, w, K=7, kernel=2)
    assert len(np.unique(lane[3]["nacc"])) > 1 or alg == Algorithms.RK4
    for a, b in zip(lane[:3], mfma[:3]):
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_ring_adjoint_checkpoint_interval_and_overflow(oracle):
    spec = _ring(Algorithms.RKCK45, ns=6)
    rs = W.Stream(W.SEED, 301)
    w = rs.normal(3 * spec["ns"] * spec["ups"]).reshape(3, -1)
    print("ring adjoint K sweep", flush=True)
    base = _adjoint(spec, w, K=1)
    for K in (5, 64, 100000):
        other = _adjoint(spec, w, K=K)
        for a, b in zip(base[:3], other[:3]):
            assert np.array_equal(a, b), K
    n = _steps(spec)
    s = _solver(spec)
    try:
        lam0, gU, gS = s.adjoint(w, 4, 2, n + 2)       # capacity far too small
        s.SynchroniseFromDeviceToHost(PB.Status)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        st = s.host_array(PB.Status).copy()
    finally:
        s.close()
    assert (st & 16).all() and np.isnan(gS).all() and np.isnan(lam0).all()


@pytest.mark.parametrize("form", [(0, 0), (0, 2), (1, 2), (1, 0)], ids=["full", "banded", "circular_banded", "full_circular"])
def test_linear_ring_adjoint_all_storage_forms(oracle, form):
    import pyoracle
    cci, cbw = form
    ns, ups = 5, 32
    rs = W.Stream(W.SEED, 310 + 2 * cci + cbw)
    s = W.new_pb_spec("pb_linear", Algorithms.RKCK45, ns, ups, cbw=cbw, cci=cci)
    Wd = 2 * cbw + 1 if cbw > 0 else ups
    s["cMat"] = rs.normal(W.coupling_storage_size(ups, cbw, cci)) * (0.5 / np.sqrt(Wd))
    s["uPAR"][0] = rs.uniform(ns * ups, -1.0, 0.3)
    s["X"][0] = rs.uniform(ns * ups, -1.0, 1.0)
    s["strength"][0] = rs.uniform(ns, 0.5, 1.5)
    s["TD"][1] = 2.0
    s["rtol"] = s["atol"] = 1e-9
    w = rs.normal(ns * ups).reshape(1, -1)
    print(f"linear adjoint form={form}", flush=True)
    lam0, gU, gS, _ = _adjoint(s, w, K=8)
    jac = pyoracle.sensitivity_pb("pb_linear", s)
    _, gx0, ga0 = _expected(s, jac, w)
    assert np.allclose(lam0[0, ::ups], gx0[:, 0], rtol=1e-9, atol=1e-12)
    assert np.allclose(gU[0, ::ups], ga0[:, 0], rtol=1e-9, atol=1e-12)


def test_ring_adjoint_rk4_finite_differences(oracle):
    """Fixed-step RK4: the adjoint is the derivative of the map itself; central
    differences (h = 1e-6) of GPU forward solves agree to 1e-6 relative."""
    spec = _ring(Algorithms.RK4, ns=4)
    rs = W.Stream(W.SEED, 302)
    w = rs.normal(3 * spec["ns"] * spec["ups"]).reshape(3, -1)
    print("ring adjoint FD", flush=True)
    _, _, gS, _ = _adjoint(spec, w, K=32)

    def loss(B):
        sp = dict(spec)
        sp["sPAR"] = B.reshape(1, -1)
        s = _solver(sp)
        try:
            s.solve_blocking()
            x = s.host_array(PB.ActualState).copy()
        finally:
            s.close()
        return (w * x).reshape(3, spec["ns"], spec["ups"]).sum(axis=(0, 2))
    B = spec["sPAR"][0].copy()
    e = 1e-6
    fd = (loss(B + e) - loss(B - e)) / (2 * e)     # each system depends only on its own B
    assert np.allclose(gS[0], fd, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("alg", [Algorithms.RK4, Algorithms.RKCK45], ids=["rk4", "rkck45"])
def test_torch_autograd_ring_period_loss(oracle, alg):
    """Loss per system: mean over units of the mean of x^2 over the last
    forcing period (two segments, quadrature reset between them); gradient
    with respect to B through torch.autograd against the forward sensitivities."""
    import torch
    import pyoracle
    from adjoint_torch import RingPeriodLoss
    spec = _ring(alg, ns=8, periods=2)
    B = torch.tensor(spec["sPAR"][0].copy(), requires_grad=True)
    print(f"torch ring loss alg={int(alg)}", flush=True)
    L = RingPeriodLoss.apply(B, spec, 16)
    L.sum().backward()
    # reference: the same two segments with the oracle's sensitivities
    TWO_PI = 2.0 * W.PI
    H = spec["TD"][1, 0]
    s1 = dict(spec)
    s1["TD"] = np.stack([np.zeros(spec["ns"]), np.full(spec["ns"], H - TWO_PI)])
    s1["X"] = spec["X"].copy()
    s1["X"][2] = 0.0
    r1 = pyoracle.run_pb("ring_duffing_q", s1)
    j1 = pyoracle.sensitivity_pb("ring_duffing_q", s1)
    s2 = dict(spec)
    s2["TD"] = np.stack([np.full(spec["ns"], H - TWO_PI), np.full(spec["ns"], H)])
    s2["X"] = r1["X"].copy()
    s2["X"][2] = 0.0
    r2 = pyoracle.run_pb("ring_duffing_q", s2)
    ups = spec["ups"]
    want = r2["X"][2].reshape(spec["ns"], ups).mean(axis=1) / TWO_PI
    assert np.allclose(L.detach().numpy(), want, rtol=1e-13, atol=0)
    # dL/dB = direct part of segment 2 + (dL/dx_mid) . (dx_mid/dB of segment 1); the
    # first needs the segment-2 sensitivity to B, the second the full Jacobian of
    # segment 2 -- finite differences of the oracle over B give both at once
    e = 1e-6

    def total(Bv):
        a = dict(s1); a["sPAR"] = Bv.reshape(1, -1)
        m = pyoracle.run_pb("ring_duffing_q", a)
        b = dict(s2); b["sPAR"] = Bv.reshape(1, -1); b["X"] = m["X"].copy(); b["X"][2] = 0.0
        return pyoracle.run_pb("ring_duffing_q", b)["X"][2].reshape(spec["ns"], ups).mean(axis=1) / TWO_PI
    B0 = spec["sPAR"][0].copy()
    fd = (total(B0 + e) - total(B0 - e)) / (2 * e)
    tol = 1e-6 if alg == Algorithms.RK4 else 1e-4
    assert np.allclose(B.grad.numpy(), fd, rtol=tol, atol=1e-9)
    # and the segment-1 part exactly against the sensitivities
    assert j1.shape[1] == 5


def test_ring_adjoint_at_256_units(oracle):
    """The adjoint at the per-block configuration's size (256 units per
    system, dense coupling on the MFMA record and reverse kernels, RKCK45 at
    1e-9) against the frozen-step forward sensitivities on a few systems."""
    import pyoracle
    spec = W.config4_ring(ns=8, ups=256, periods=1, system="ring_duffing_q")
    rs = W.Stream(W.SEED, 305)
    w = rs.normal(3 * spec["ns"] * spec["ups"]).reshape(3, -1)
    print("ring adjoint 8 x 256 units, MFMA", flush=True)
    lam0, gU, gS, out = _adjoint(spec, w, K=32, kernel=2)
    ref = pyoracle.run_pb("ring_duffing_q", spec)
    assert np.array_equal(out["X"], ref["X"]) and np.array_equal(out["nacc"], ref["nacc"])
    jac = pyoracle.sensitivity_pb("ring_duffing_q", spec)
    gB, gx0, gk0 = _expected(spec, jac, w)
    ups = spec["ups"]
    assert np.allclose(gS[0], gB[:, 0], rtol=1e-8, atol=1e-12)
    for c in range(3):
        assert np.allclose(lam0[c, ::ups], gx0[:, c], rtol=1e-8, atol=1e-12)
    assert np.allclose(gU[0, ::ups], gk0[:, 0], rtol=1e-8, atol=1e-12)
