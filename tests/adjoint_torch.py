"""A thin torch.autograd.Function over the library's discrete adjoint (test
harness only; the product exposes the adjoint through the C ABI).

DuffingPeriodLoss(k, B, x0, v0) -> per-instance loss L_i = mean of x^2 over the
last forcing period of a horizon of `periods` periods, computed as two Solve
segments of the duffing_q system (x, v, q' = x^2): [0, H - 2pi] and
[H - 2pi, H] with q reset to 0, L = q(H) / 2pi.  backward() runs the adjoint
of segment 2 (seed dL/dq = 1/2pi), then of segment 1 with the q component of
the adjoint cut (q is reset between the segments)."""
import numpy as np
import torch

from mpgos import Algorithms, ListOfVariables as V, ProblemSolver
from mpgos import workloads as W

TWO_PI = 2.0 * W.PI


def _solver(nt, algorithm, k, B, x0, v0, t0, t1, q0, opts):
    s = ProblemSolver("duffing_q", nt, algorithm)
    spec = W.new_spec("duffing_q", algorithm, nt)
    spec["cPAR"][0], spec["cPAR"][1] = k, B
    spec["X"][0], spec["X"][1], spec["X"][2] = x0, v0, q0
    spec["TD"][0], spec["TD"][1] = t0, t1
    spec.update(opts)
    spec["options"] = dict(opts.get("options", {}))
    W.load(s, spec)
    return s


def _capacity(solver, K):
    # a plain Solve on a throw-away copy would be exact; the step count of the
    # recorded pass is bounded by the horizon over the minimum step actually
    # used, so run once and size from the accepted counts
    solver.solve_blocking()
    n = int(solver.host_array(V.AcceptedSteps).max())
    return n // K + 2


class DuffingPeriodLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k, B, x0, v0, algorithm, periods, K, opts):
        nt = k.numel()
        H = periods * TWO_PI
        kk, bb, xx, vv = (t.detach().cpu().numpy().astype(np.float64) for t in (k, B, x0, v0))
        s1 = _solver(nt, algorithm, kk, bb, xx, vv, 0.0, H - TWO_PI, 0.0, opts)
        cap1 = _capacity(s1, K)
        W.load(s1, _reload(s1, kk, bb, xx, vv, 0.0, H - TWO_PI, opts, algorithm))
        s1.adjoint_record(K, cap1)
        s1.SynchroniseFromDeviceToHost(V.All)
        s1.InsertSynchronisationPoint()
        s1.SynchroniseSolver()
        xm = s1.host_array(V.ActualState).copy()
        s2 = _solver(nt, algorithm, kk, bb, xm[0], xm[1], H - TWO_PI, H, 0.0, opts)
        cap2 = _capacity(s2, K)
        W.load(s2, _reload(s2, kk, bb, xm[0], xm[1], H - TWO_PI, H, opts, algorithm))
        s2.adjoint_record(K, cap2)
        s2.SynchroniseFromDeviceToHost(V.All)
        s2.InsertSynchronisationPoint()
        s2.SynchroniseSolver()
        loss = s2.host_array(V.ActualState)[2].copy() / TWO_PI
        ctx.solvers = (s1, s2)
        return torch.from_numpy(loss)

    @staticmethod
    def backward(ctx, g):
        s1, s2 = ctx.solvers
        g = g.detach().cpu().numpy().astype(np.float64)
        lam = np.zeros((3, g.size))
        lam[2] = g / TWO_PI
        s2.host_array(V.AdjointState)[...] = lam
        s2.SynchroniseFromHostToDevice(V.AdjointState)
        s2.adjoint_reverse()
        s2.SynchroniseFromDeviceToHost(V.AdjointState)
        s2.SynchroniseFromDeviceToHost(V.ParameterGradient)
        s2.InsertSynchronisationPoint()
        s2.SynchroniseSolver()
        lam_mid = s2.host_array(V.AdjointState).copy()
        gp = s2.host_array(V.ParameterGradient).copy()
        lam_mid[2] = 0.0                                   # q restarts at 0 in segment 2
        s1.host_array(V.AdjointState)[...] = lam_mid
        s1.SynchroniseFromHostToDevice(V.AdjointState)
        s1.adjoint_reverse()
        s1.SynchroniseFromDeviceToHost(V.AdjointState)
        s1.SynchroniseFromDeviceToHost(V.ParameterGradient)
        s1.InsertSynchronisationPoint()
        s1.SynchroniseSolver()
        lam0 = s1.host_array(V.AdjointState).copy()
        gp = gp + s1.host_array(V.ParameterGradient)
        s1.close()
        s2.close()
        t = torch.from_numpy
        return t(gp[0]), t(gp[1]), t(lam0[0]), t(lam0[1]), None, None, None, None


def _reload(solver, k, B, x0, v0, t0, t1, opts, algorithm):
    spec = W.new_spec("duffing_q", algorithm, solver.NT)
    spec["cPAR"][0], spec["cPAR"][1] = k, B
    spec["X"][0], spec["X"][1], spec["X"][2] = x0, v0, 0.0
    spec["TD"][0], spec["TD"][1] = t0, t1
    spec.update({k_: v for k_, v in opts.items() if k_ != "options"})
    spec["options"] = dict(opts.get("options", {}))
    return spec


def duffing_period_loss(k, B, x0, v0, algorithm=Algorithms.RK4, periods=64, K=32, opts=None):
    return DuffingPeriodLoss.apply(k, B, x0, v0, algorithm, periods, K, opts or {})


class RingPeriodLoss(torch.autograd.Function):
    """Per-system loss of a per-block ring with quadrature state
    (ring_duffing_q): L_s = mean over units of the mean of x^2 over the last
    forcing period, as two segments [0, H - 2pi] and [H - 2pi, H] with q reset
    to 0.  Input: B (one per system).  backward() runs the per-block adjoint of
    segment 2 (seed dL/dq = 1/(2pi U)), cuts the q component, then segment 1,
    and returns dL/dB summed over both segments."""

    @staticmethod
    def forward(ctx, B, spec, K):
        from mpgos.perblock import PB, PBProblemSolver
        ns, ups = spec["ns"], spec["ups"]
        H = spec["TD"][1, 0]
        b = B.detach().cpu().numpy().astype(np.float64)

        def run(x0, t0, t1):
            sp = dict(spec)
            sp["sPAR"] = b.reshape(1, -1)
            sp["X"] = x0.copy()
            sp["X"][2] = 0.0
            sp["TD"] = np.stack([np.full(ns, t0), np.full(ns, t1)])
            s = PBProblemSolver("ring_duffing_q", ns, ups, spec["algorithm"])
            W.load_pb(s, sp)
            s.solve_blocking()                      # sizes the record buffers
            n = int(s.host_array(PB.AcceptedSteps).max())
            W.load_pb(s, sp)
            s.adjoint_record(K, n // K + 2, n + 2)
            s.SynchroniseFromDeviceToHost(PB.All)
            s.InsertSynchronisationPoint()
            s.SynchroniseSolver()
            return s, s.host_array(PB.ActualState).copy()
        s1, xm = run(spec["X"], 0.0, H - TWO_PI)
        s2, xe = run(xm, H - TWO_PI, H)
        ctx.solvers = (s1, s2)
        ctx.shape = (ns, ups)
        return torch.from_numpy(xe[2].reshape(ns, ups).mean(axis=1) / TWO_PI)

    @staticmethod
    def backward(ctx, g):
        from mpgos.perblock import PB
        s1, s2 = ctx.solvers
        ns, ups = ctx.shape
        g = g.detach().cpu().numpy().astype(np.float64)

        def reverse(s, lam):
            s.host_array(PB.AdjointState)[...] = lam
            s.SynchroniseFromHostToDevice(PB.AdjointState)
            s.adjoint_reverse()
            for v in (PB.AdjointState, PB.SystemParameterGradient):
                s.SynchroniseFromDeviceToHost(v)
            s.InsertSynchronisationPoint()
            s.SynchroniseSolver()
            return s.host_array(PB.AdjointState).copy(), s.host_array(PB.SystemParameterGradient).copy()
        lam = np.zeros((3, ns * ups))
        lam[2] = np.repeat(g / (TWO_PI * ups), ups)
        lam_mid, gB2 = reverse(s2, lam)
        lam_mid[2] = 0.0                            # q restarts at 0 in segment 2
        _, gB1 = reverse(s1, lam_mid)
        s1.close()
        s2.close()
        return torch.from_numpy(gB1[0] + gB2[0]), None, None
