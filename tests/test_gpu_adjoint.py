"""Discrete adjoint on the GPU (FP64) against the oracle's forward
sensitivities on the same frozen step sequence, and against central finite
differences where the map is smooth (fixed steps)."""
import numpy as np
import pytest

import scenarios as S
from mpgos import Algorithms, ListOfVariables as V, ProblemSolver
from mpgos import workloads as W

pytestmark = pytest.mark.gpu


def gpu_adjoint(spec, lam_T, K, cap=None):
    s = ProblemSolver(spec["system"], spec["nt"], spec["algorithm"])
    try:
        W.load(s, spec)
        if cap is None:
            s.solve_blocking()
            cap = int(s.host_array(V.AcceptedSteps).max()) // K + 2
            W.load(s, spec)
        lam0, gp = s.adjoint(lam_T, K, cap)
        s.SynchroniseFromDeviceToHost(V.All)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        out = W.read(s)
    finally:
        s.close()
    return lam0, gp, out


def reference_gradient(oracle, spec, lam_T):
    o = oracle.sensitivity(spec["system"], spec)
    jac = o["jac"]                                      # (SD, SD + NCP, NT)
    g = np.einsum("in,ijn->jn", lam_T, jac)
    return g[:spec["sd"]], g[spec["sd"]:], o


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("algorithm", [Algorithms.RK4, Algorithms.RKCK45])
def test_duffing_q_adjoint_matches_forward_sensitivity(gpu, oracle, algorithm):
    spec = W.config3_duffing_adjoint(nt=512, algorithm=algorithm)
    spec["TD"][1] = 4 * (2 * W.PI)
    rng = np.random.default_rng(3)
    lam_T = rng.normal(size=(3, 512))
    lam0, gp, out = gpu_adjoint(spec, lam_T, K=16)
    rx, rp, ref = reference_gradient(oracle, spec, lam_T)
    assert np.array_equal(out["X"].view(np.uint64), ref["X"].view(np.uint64))   # recorded forward = oracle
    assert rel(lam0, rx) < 1e-9, rel(lam0, rx)
    assert rel(gp, rp) < 1e-9, rel(gp, rp)


@pytest.mark.parametrize("algorithm", [Algorithms.RK4, Algorithms.RKCK45], ids=["rk4", "rkck45"])
def test_adjoint_independent_of_checkpoint_interval(gpu, oracle, algorithm):
    """Checkpoints every 1, 7 or 64 accepted steps (segments recomputed from
    them) give the same gradients bit for bit."""
    spec = W.config3_duffing_adjoint(nt=256, algorithm=algorithm)
    spec["TD"][1] = 2 * (2 * W.PI)
    lam_T = np.ones((3, 256))
    res = [gpu_adjoint(spec, lam_T, K=K)[:2] for K in (1, 7, 64)]
    for r in res[1:]:
        assert np.array_equal(r[0].view(np.uint64), res[0][0].view(np.uint64))
        assert np.array_equal(r[1].view(np.uint64), res[0][1].view(np.uint64))


def test_lorenz_adjoint(gpu, oracle):
    spec = W.config1_lorenz(nt=512)
    spec["TD"][1] = 1.0
    lam_T = np.random.default_rng(5).normal(size=(3, 512))
    lam0, gp, _ = gpu_adjoint(spec, lam_T, K=32)
    rx, rp, _ = reference_gradient(oracle, spec, lam_T)
    assert rel(lam0, rx) < 1e-9 and rel(gp, rp) < 1e-9


def test_affine_adjoint_closed_form(gpu, oracle):
    A = np.array([[-0.5, 1.0], [-1.0, -0.1]])
    spec = S.affine(nt=64, a00=A[0, 0], a01=A[0, 1], a10=A[1, 0], a11=A[1, 1], b0=0.3, b1=-0.2, x0=(1.0, 0.5),
                    td=(0.0, 2.0), dt0=0.125, algorithm=Algorithms.RK4)
    lam_T = np.tile(np.array([[1.0], [-2.0]]), (1, 64))
    lam0, gp, _ = gpu_adjoint(spec, lam_T, K=4)
    hA = 0.125 * A
    M = np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    expect = np.linalg.matrix_power(M, 16).T @ lam_T[:, 0]
    assert np.allclose(lam0[:, 0], expect, rtol=1e-13)


def test_finite_differences_rk4(gpu, oracle):
    # central differences (h = 1e-6, relative) of L = x(T) . lam on the RK4 map
    spec = W.config3_duffing_adjoint(nt=64, algorithm=Algorithms.RK4)
    spec["TD"][1] = 2 * (2 * W.PI)
    lam_T = np.array([[1.0], [0.5], [1.0]]) * np.ones((3, 64))
    lam0, gp, out = gpu_adjoint(spec, lam_T, K=32)
    L = lambda o: (lam_T * o["X"]).sum(0)
    for j, (key, row) in enumerate((("cPAR", 0), ("cPAR", 1), ("X", 0), ("X", 1))):
        hstep = 1e-6 * np.abs(spec[key][row]).max()
        sp, sm = dict(spec), dict(spec)
        sp[key] = spec[key].copy(); sm[key] = spec[key].copy()
        sp[key][row] += hstep
        sm[key][row] -= hstep
        fd = (L(oracle.run("duffing_q", sp)) - L(oracle.run("duffing_q", sm))) / (2 * hstep)
        ad = gp[row] if key == "cPAR" else lam0[row]
        assert rel(ad, fd) < 1e-6, (key, row, rel(ad, fd))


def test_checkpoint_overflow_is_reported(gpu, oracle):
    spec = W.config3_duffing_adjoint(nt=64, algorithm=Algorithms.RK4)
    spec["TD"][1] = 2 * W.PI
    lam0, gp, out = gpu_adjoint(spec, np.ones((3, 64)), K=4, cap=3)
    assert (out["status"] & 16).all() and np.isnan(gp).all()


@pytest.mark.parametrize("algorithm", [Algorithms.RK4, Algorithms.RKCK45])
def test_torch_autograd_two_segment_loss(gpu, oracle, algorithm):
    # the adjoint-configuration loss: mean of x^2 over the last forcing period
    import torch
    from adjoint_torch import duffing_period_loss
    nt, periods = 128, 4
    spec = W.config3_duffing_adjoint(nt=nt, algorithm=algorithm, periods=periods)
    opts = {} if algorithm == Algorithms.RK4 else {"rtol": 1e-9, "atol": 1e-9}
    if algorithm == Algorithms.RK4:
        opts["options"] = {"dt0": 2 * W.PI / 500}
    k = torch.tensor(spec["cPAR"][0], requires_grad=True)
    B = torch.tensor(spec["cPAR"][1], requires_grad=True)
    x0 = torch.tensor(spec["X"][0], requires_grad=True)
    v0 = torch.tensor(spec["X"][1], requires_grad=True)
    loss = duffing_period_loss(k, B, x0, v0, algorithm=algorithm, periods=periods, K=32, opts=opts)
    loss.sum().backward()
    # reference: compose the oracle's sensitivities of the two segments
    two_pi = 2 * W.PI
    s1 = dict(spec)
    o1 = oracle.sensitivity("duffing_q", s1)
    s2 = W.new_spec("duffing_q", algorithm, nt)
    s2["cPAR"] = spec["cPAR"].copy()
    s2["X"][0], s2["X"][1] = o1["X"][0], o1["X"][1]
    s2["TD"][0], s2["TD"][1] = (periods - 1) * two_pi, periods * two_pi
    s2["T"][:] = s2["TD"][0]
    for key in ("rtol", "atol"):
        s2[key] = spec[key]
    s2["options"] = dict(spec["options"])
    o2 = oracle.sensitivity("duffing_q", s2)
    assert np.allclose(loss.detach().numpy(), o2["X"][2] / two_pi, rtol=1e-14, atol=0)
    J1, J2 = o1["jac"], o2["jac"]                  # (3, 5, nt): z = (x, v, q, k, B)
    dq_dxv = J2[2, :2]                             # (2, nt)
    dL_dp = (J2[2, 3:] + np.einsum("an,apn->pn", dq_dxv, J1[:2, 3:])) / two_pi
    dL_dx0 = np.einsum("an,apn->pn", dq_dxv, J1[:2, :2]) / two_pi
    assert rel(k.grad.numpy(), dL_dp[0]) < 1e-6 and rel(B.grad.numpy(), dL_dp[1]) < 1e-6
    assert rel(x0.grad.numpy(), dL_dx0[0]) < 1e-6 and rel(v0.grad.numpy(), dL_dx0[1]) < 1e-6


def test_checkpoint_interval_autotuner(gpu, oracle):
    """The autotuner times record + reverse for each candidate K, leaves the
    solver's inputs as they were, and the gradients with its K equal those
    with K = 1 bit for bit."""
    from mpgos.tuning import autotune_adjoint
    spec = W.config3_duffing_adjoint(nt=256, algorithm=Algorithms.RKCK45)
    spec["TD"][1] = 2 * (2 * W.PI)
    lam_T = np.random.default_rng(5).normal(size=(3, 256))
    base = gpu_adjoint(spec, lam_T, K=1)
    s = ProblemSolver(spec["system"], spec["nt"], spec["algorithm"])
    try:
        W.load(s, spec)
        x0 = spec["X"].copy()
        nmax = int(base[2]["nacc"].max())
        K, table = autotune_adjoint(s, lam_T, nmax, candidates=(4, 16, 64))
        assert K in (4, 16, 64) and len(table) == 3
        assert all(r["record_ms"] > 0 and r["reverse_ms"] > 0 for r in table)
        s.SynchroniseFromDeviceToHost(V.ActualState)
        s.InsertSynchronisationPoint()
        s.SynchroniseSolver()
        assert np.array_equal(s.host_array(V.ActualState), x0)      # inputs restored
        lam0, gp = s.adjoint(lam_T, K, nmax // K + 2)
    finally:
        s.close()
    assert np.array_equal(lam0.view(np.uint64), base[0].view(np.uint64))
    assert np.array_equal(gp.view(np.uint64), base[1].view(np.uint64))


def _ck45_replay(x0, k, B, ts, hs):
    """Cash-Karp fifth-order solution of the Duffing+q system on given
    accepted steps (t_n, h_n), vectorised over columns; zero steps leave a
    column unchanged.  Used only to form finite differences on a frozen step
    sequence."""
    def f(y, t):
        return np.stack([y[1], -k * y[1] + B * np.cos(t) - (y[0] ** 3 + y[0]), y[0] ** 2])
    x = x0.copy()
    for t, h in zip(ts, hs):
        k1 = f(x, t)
        k2 = f(x + h * (k1 / 5), t + h / 5)
        k3 = f(x + h * (3 / 40 * k1 + 9 / 40 * k2), t + 3 / 10 * h)
        k4 = f(x + h * (3 / 10 * k1 - 9 / 10 * k2 + 6 / 5 * k3), t + 3 / 5 * h)
        k5 = f(x + h * (-11 / 54 * k1 + 5 / 2 * k2 - 70 / 27 * k3 + 35 / 27 * k4), t + h)
        k6 = f(x + h * (1631 / 55296 * k1 + 175 / 512 * k2 + 575 / 13824 * k3 + 44275 / 110592 * k4 + 253 / 4096 * k5),
               t + 7 / 8 * h)
        x = x + h * (37 / 378 * k1 + 250 / 621 * k3 + 125 / 594 * k4 + 512 / 1771 * k6)
    return x


def test_finite_differences_rkck45_frozen_steps(gpu, oracle):
    """Central differences (h = 1e-6, relative) of L = lam . x(T) on the
    accepted-step sequence of the nominal adaptive run, replayed with the
    step sizes held fixed: the derivative the adjoint of the adaptive Solve
    computes (the controller is not differentiated)."""
    nt = 32
    spec = W.config3_duffing_adjoint(nt=nt, algorithm=Algorithms.RKCK45)
    spec["TD"][1] = 2 * (2 * W.PI)
    lam_T = np.array([[1.0], [0.5], [1.0]]) * np.ones((3, nt))
    lam0, gp, out = gpu_adjoint(spec, lam_T, K=16)
    # the accepted steps of every instance from the oracle's per-attempt trace
    seqs = []
    for i in range(nt):
        tr = oracle.trace("duffing_q", spec, tid=i)
        a = tr["accepted"] == 1
        seqs.append((tr["t"][a], tr["h"][a]))
        assert len(seqs[-1][0]) == out["nacc"][i]
    N = max(len(t) for t, _ in seqs)
    ts = np.zeros((N, nt))
    hs = np.zeros((N, nt))
    for i, (t, h) in enumerate(seqs):
        ts[:len(t), i] = t
        hs[:len(h), i] = h
    k, B = spec["cPAR"][0], spec["cPAR"][1]
    x0 = spec["X"].copy()
    nominal = _ck45_replay(x0, k, B, ts, hs)
    assert np.allclose(nominal, out["X"], rtol=1e-12, atol=1e-13)   # the replay is the recorded Solve
    L = lambda x: (lam_T * x).sum(0)
    for name, ad in (("k", gp[0]), ("B", gp[1]), ("x0", lam0[0]), ("v0", lam0[1])):
        p = {"k": k, "B": B, "x0": x0[0], "v0": x0[1]}[name]
        hstep = 1e-6 * np.abs(p).max()

        def run(sign):
            kk, BB, xx = k.copy(), B.copy(), x0.copy()
            if name == "k": kk = kk + sign * hstep
            elif name == "B": BB = BB + sign * hstep
            elif name == "x0": xx[0] = xx[0] + sign * hstep
            else: xx[1] = xx[1] + sign * hstep
            return L(_ck45_replay(xx, kk, BB, ts, hs))
        fd = (run(1) - run(-1)) / (2 * hstep)
        assert rel(ad, fd) < 1e-6, (name, rel(ad, fd))


@pytest.mark.parametrize("algorithm", [Algorithms.RK4, Algorithms.RKCK45], ids=["rk4", "rkck45"])
def test_adjoint_of_a_system_with_accessories(gpu, oracle, algorithm):
    """Duffing with its max |x| accessory (updated after every accepted step):
    the checkpoints carry the accessories, the recorded Solve's outputs equal
    the oracle's bit for bit and the gradients its sensitivities."""
    spec = W.config0_duffing_rk4(nt=256)
    spec["algorithm"] = int(algorithm)
    spec["TD"][1] = 3 * (2 * W.PI)
    if algorithm == Algorithms.RKCK45:
        spec["rtol"] = spec["atol"] = 1e-9
    lam_T = np.random.default_rng(11).normal(size=(2, 256))
    lam0, gp, out = gpu_adjoint(spec, lam_T, K=5)
    rx, rp, ref = reference_gradient(oracle, spec, lam_T)
    assert np.array_equal(out["ACC"].view(np.uint64), ref["ACC"].view(np.uint64))
    assert rel(lam0, rx) < 1e-9 and rel(gp, rp) < 1e-9
