"""Per-block coupled systems on the GPU, bit-for-bit against the oracle:
the reference tutorial's double-well ring (events, accessories, dense output,
several systems per block), linear rings in every coupling storage form with
the matrix in L2 or staged in LDS, partial activity, and the dense forced
Duffing ring of 256 units and of 96, 160 and 224 units (the dense-contraction
form with the tiled matrix wholly or partly in LDS).  Calls go through libmpgos's per-block C ABI."""
import numpy as np
import pytest

from mpgos import workloads as W
from mpgos.enums import Algorithms
from mpgos._lib import MPGOSError
from mpgos.perblock import PB, PBOption, PBProblemSolver

pytestmark = pytest.mark.gpu

KEYS = ("X", "T", "TD", "uACC", "uACCi", "sACC", "sACCi", "DOidx", "nacc", "nrej", "status", "DOt", "DOx")


def run_pb_gpu(spec, spb=None, shared_matrix=None, shared_globals=None, solves=1, kernel=None):
    s = PBProblemSolver(spec["system"], spec["ns"], spec["ups"], spec["algorithm"], spb=spb or spec.get("spb", 1),
                        cbw=spec["cbw"], cci=spec["cci"], ndo=spec["ndo"])
    try:
        W.load_pb(s, spec)
        if kernel is not None:
            s.SolverOption(PBOption.CouplingKernel, kernel)
        if shared_matrix is not None:
            s.SolverOption(PBOption.SharedCouplingMatrices, shared_matrix)
        if shared_globals is not None:
            s.SolverOption(PBOption.SharedGlobalVariables, shared_globals)
        outs = []
        for _ in range(solves):
            s.solve_blocking()
            outs.append(W.read_pb(s))
        return outs if solves > 1 else outs[0]
    finally:
        s.close()


def _mask_dense(out):
    out = dict(out)
    if "DOt" in out and out["DOt"].size:
        ndo, ns = out["DOt"].shape
        valid = np.arange(ndo)[:, None] < np.minimum(out["DOidx"], ndo)[None, :]
        out["DOt"] = np.where(valid, out["DOt"], 0.0)
        ups = out["DOx"].shape[2] // ns
        vu = np.repeat(valid, ups, axis=1)
        out["DOx"] = np.where(vu[:, None, :], out["DOx"], 0.0)
    return out


def assert_pb_bitwise(gpu, ref, label=""):
    gpu, ref = _mask_dense(gpu), _mask_dense(ref)
    for k in KEYS:
        if k not in gpu or k not in ref:
            continue
        a, b = np.asarray(gpu[k]), np.asarray(ref[k])
        assert a.shape == b.shape, f"{label} {k}: shape {a.shape} vs {b.shape}"
        same = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b)) if a.dtype == np.float64 else a == b
        if not same.all():
            bad = np.argwhere(~same)
            i = tuple(bad[0])
            raise AssertionError(f"{label} {k}: {len(bad)} of {a.size} differ; first at {i}: gpu={a[i]!r} oracle={b[i]!r}")


@pytest.mark.parametrize("alg", [Algorithms.RKCK45, Algorithms.RK4], ids=["rkck45", "rk4"])
def test_tutorial_ring_dw(oracle, alg):
    import pyoracle
    spec = W.tutorial_ring_dw(algorithm=alg)
    if alg == Algorithms.RK4:
        spec["options"]["dt0"] = 1e-3
    print(f"tutorial ring_dw alg={int(alg)}", flush=True)
    ref = pyoracle.run_pb("ring_dw", spec)
    for spb in (3, 1, 4):
        got = run_pb_gpu(spec, spb=spb)
        assert_pb_bitwise(got, ref, f"spb={spb}")
    assert (ref["uACCi"].reshape(-1) <= 2).all() and (ref["nacc"] > 10).all()


FORMS = [(0, 0), (0, 2), (1, 2), (1, 0)]


@pytest.mark.parametrize("form", FORMS, ids=["full", "banded", "circular_banded", "full_circular"])
@pytest.mark.parametrize("alg", [Algorithms.RKCK45, Algorithms.RK4], ids=["rkck45", "rk4"])
def test_linear_ring_forms(oracle, form, alg):
    import pyoracle
    cci, cbw = form
    ns, ups = 37, 32
    rs = W.Stream(W.SEED, 200 + 2 * cci + cbw)
    s = W.new_pb_spec("pb_linear", alg, ns, ups, cbw=cbw, cci=cci, ndo=8)
    Wd = 2 * cbw + 1 if cbw > 0 else ups
    s["cMat"] = rs.normal(W.coupling_storage_size(ups, cbw, cci)) * (0.5 / np.sqrt(Wd))
    s["uPAR"][0] = rs.uniform(ns * ups, -1.0, 0.3)
    s["X"][0] = rs.uniform(ns * ups, -1.0, 1.0)
    s["strength"][0] = rs.uniform(ns, 0.5, 1.5)
    s["TD"][1] = rs.uniform(ns, 1.0, 3.0)
    s["rtol"] = s["atol"] = 1e-9
    s["options"].update(dt0=1.0 / 128.0, do_freq=3)
    s["n_active"] = ns - 5
    print(f"linear ring form={form} alg={int(alg)}", flush=True)
    ref = pyoracle.run_pb("pb_linear", s)
    # inactive systems are untouched
    assert np.array_equal(ref["X"][:, (ns - 5) * ups:], s["X"][:, (ns - 5) * ups:])
    for spb, shm in ((1, 0), (2, 1), (5, 0)):
        got = run_pb_gpu(s, spb=spb, shared_matrix=shm, kernel=1)
        assert_pb_bitwise(got, ref, f"lane kernel spb={spb} shared_matrix={shm}")
    if form == (0, 0):
        for shm in (0, 1):
            got = run_pb_gpu(s, shared_matrix=shm, kernel=2)
            assert_pb_bitwise(got, ref, f"mfma kernel shared_matrix={shm}")
    else:
        with pytest.raises(MPGOSError, match="dense-contraction"):
            run_pb_gpu(s, kernel=2)


def test_ring_dw_full_matrix_mfma(oracle):
    """The tutorial's double-well ring with 32 units and its nearest-neighbour
    coupling stored as a full matrix: events, accessories, user termination
    and dense output through the dense-contraction kernel."""
    import pyoracle
    ns, ups = 11, 32
    base = W.tutorial_ring_dw(ns=25, ups=ups)
    s = W.new_pb_spec("ring_dw", Algorithms.RKCK45, ns, ups, ndo=64)
    for k in ("rtol", "atol", "edir", "options", "gPAR", "cIndex"):
        s[k] = base[k]
    for key in ("TD", "T", "sPAR", "sACC", "sACCi", "strength", "DOidx"):
        s[key] = np.ascontiguousarray(base[key][..., :ns])
    for key in ("X", "uPAR", "uACC", "uACCi"):
        s[key] = np.ascontiguousarray(base[key][:, :ns * ups])
    ring = np.zeros((ups, ups))
    for r in range(ups):
        ring[r, (r + 1) % ups] = ring[r, (r - 1) % ups] = 1.0
    s["cMat"] = np.concatenate([W.store_coupling(ring, 0, 0)] * 2)
    print("ring_dw full matrix 32 units", flush=True)
    ref = pyoracle.run_pb("ring_dw", s)
    for kernel in (1, 2):
        got = run_pb_gpu(s, kernel=kernel)
        assert_pb_bitwise(got, ref, f"kernel={kernel}")
    assert (ref["uACCi"] == 2).any() and (ref["T"] < s["TD"][1]).all()


@pytest.mark.parametrize("kernel", [1, 2], ids=["lane", "mfma"])
@pytest.mark.parametrize("alg", [Algorithms.RKCK45, Algorithms.RK4], ids=["rkck45", "rk4"])
def test_dense_duffing_ring_256(oracle, alg, kernel):
    import pyoracle
    spec = W.config4_ring(ns=27, ups=256, periods=1, algorithm=alg)
    if alg == Algorithms.RK4:
        spec["options"]["dt0"] = 2.0 * W.PI / 400.0
    spec["n_active"] = 25
    print(f"dense ring 256 alg={int(alg)} kernel={kernel}", flush=True)
    ref = pyoracle.run_pb("ring_duffing", spec)
    for shm in (0, 1):
        got = run_pb_gpu(spec, shared_matrix=shm, kernel=kernel)
        assert_pb_bitwise(got, ref, f"shared_matrix={shm}")


@pytest.mark.parametrize("ups", [96, 160, 224])
def test_dense_ring_tile_counts(oracle, ups):
    """Unit counts that are not powers of two: 6, 10 and 14 tiles, 3, 5 and 7
    waves.  At 96 units the whole tiled matrix fits into the workgroup's LDS;
    at 160 and 224 the leading 12 and 8 column-block pairs of every tile do and
    the rest is read from global memory two groups deep.  Both units-per-lane
    layouts of the dense-contraction form against the oracle and the lane
    form, bit for bit."""
    import pyoracle
    spec = W.config4_ring(ns=11, ups=ups, periods=1)
    spec["n_active"] = 10
    print(f"dense ring ups={ups}", flush=True)
    ref = pyoracle.run_pb("ring_duffing", spec)
    assert_pb_bitwise(run_pb_gpu(spec, kernel=1), ref, f"lane ups={ups}")
    assert_pb_bitwise(run_pb_gpu(spec, kernel=2), ref, f"mfma ups={ups}")
    s = PBProblemSolver("ring_duffing", spec["ns"], ups, spec["algorithm"])
    try:
        W.load_pb(s, spec)
        s.SolverOption(PBOption.CouplingKernel, 2)
        s.SolverOption(PBOption.MfmaUnitsPerLane, 2)
        s.solve_blocking()
        assert_pb_bitwise(W.read_pb(s), ref, f"mfma 2 units per lane ups={ups}")
    finally:
        s.close()


def test_second_solve_continues(oracle):
    """A second Solve after moving the time domain continues from the stored
    state, as two oracle runs do."""
    import pyoracle
    spec = W.config4_ring(ns=8, ups=64, periods=1)
    s = PBProblemSolver("ring_duffing", 8, 64, Algorithms.RKCK45)
    try:
        W.load_pb(s, spec)
        s.solve_blocking()
        first = W.read_pb(s)
        td = s.host_array(PB.TimeDomain)
        td[0] = td[1]
        td[1] = 2.0 * td[1]
        s.SynchroniseFromHostToDevice(PB.TimeDomain)
        s.solve_blocking()
        second = W.read_pb(s)
    finally:
        s.close()
    ref1 = pyoracle.run_pb("ring_duffing", spec)
    assert_pb_bitwise(first, ref1, "first")
    spec2 = dict(spec)
    spec2["X"], spec2["T"] = ref1["X"], ref1["T"]
    spec2["TD"] = np.stack([spec["TD"][1], 2.0 * spec["TD"][1]])
    ref2 = pyoracle.run_pb("ring_duffing", spec2)
    assert_pb_bitwise(second, ref2, "second")


def test_dropin_cpp_driver_matches_oracle(oracle, tmp_path):
    """The reference tutorial's driver, compiled against the drop-in
    CoupledSystems_PerBlock_Interface.cuh, reproduces the oracle bit for bit;
    its Print files follow the reference's layout (unit scope padded to TPB)."""
    import os
    import subprocess
    import pyoracle
    exe = os.path.join(os.path.dirname(__file__), "cpp", "ring_dropin")
    assert os.path.exists(exe), "build() compiles tests/cpp/ring_dropin"
    out = tmp_path / "out.bin"
    print("ring_dropin driver", flush=True)
    r = subprocess.run([exe, str(out)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    ns, ups = 25, 14
    nu = ns * ups
    v = np.fromfile(out, dtype=np.float64)
    parts = {}
    o = 0
    for name, n in (("X", 2 * nu), ("T", ns), ("uACC", 3 * nu), ("uACCi", nu), ("sACC", 2 * ns), ("sACCi", ns), ("DOidx", ns)):
        parts[name] = v[o:o + n]
        o += n
    assert o == v.size
    ref = pyoracle.run_pb("ring_dw", W.tutorial_ring_dw())
    got = {"X": parts["X"].reshape(2, nu), "T": parts["T"], "uACC": parts["uACC"].reshape(3, nu),
           "uACCi": parts["uACCi"].reshape(1, nu).astype(np.int32), "sACC": parts["sACC"].reshape(2, ns),
           "sACCi": parts["sACCi"].reshape(1, ns).astype(np.int32), "DOidx": parts["DOidx"].astype(np.int32)}
    for k, a in got.items():
        b = np.asarray(ref[k])
        same = (a.view(np.uint64) == b.view(np.uint64)) if a.dtype == np.float64 else (a == b)
        assert same.all(), k
    rows = (tmp_path / "ActualStateInSolverObject.txt").read_text().splitlines()
    assert len(rows) == 9 * 64 and sum("PADDING" in r for r in rows) == 9 * 64 - nu
    first = [x for x in rows[0].split(",") if x.strip()]
    assert float(first[0]) == pytest.approx(ref["X"][0, 0], rel=1e-9)
    dense = (tmp_path / "DenseOutput_0.txt").read_text().splitlines()
    series = dense[dense.index("Time series:") + 2:]
    assert len(series) == ref["DOidx"][0]
    assert float(series[-1].split(",")[0]) == pytest.approx(ref["T"][0], rel=1e-9)
