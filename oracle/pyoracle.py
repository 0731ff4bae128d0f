"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(),
bench.py's cpu_baseline).  ctypes front end of oracle/liboracle.so, the FP64
OpenMP restatement of the reference's per-thread solver (pt_oracle.h).

`run(system, spec)` integrates the instances described by `spec` (a dict of
numpy arrays in the reference's SoA layout plus solver options, see
PTSpec in mpgos/workloads.py) and returns the outputs as numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MPGOS_ORACLE_LIB selects another build of the same sources (the sanitizer
# build, liboracle_san.so, in the CPU sanitizer test)
LIB = os.environ.get("MPGOS_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)


class Job(ctypes.Structure):
    _fields_ = [
        ("nt", ctypes.c_int), ("ndo", ctypes.c_int), ("algorithm", ctypes.c_int), ("n_active", ctypes.c_int),
        ("dt0", ctypes.c_double), ("dtmax", ctypes.c_double), ("dtmin", ctypes.c_double),
        ("grow", ctypes.c_double), ("shrink", ctypes.c_double), ("do_min_dt", ctypes.c_double),
        ("do_freq", ctypes.c_int),
        ("rtol", _dp), ("atol", _dp), ("etol", _dp), ("edir", _ip),
        ("evcap", ctypes.c_int),
        ("TD", _dp), ("X", _dp), ("T", _dp), ("cPAR", _dp), ("sPAR", _dp), ("sPARi", _ip),
        ("ACC", _dp), ("ACCi", _ip), ("DOidx", _ip), ("DOt", _dp), ("DOx", _dp),
        ("nacc", _lp), ("nrej", _lp), ("status", _ip), ("nev", _ip), ("evlog", _dp), ("evlogidx", _ip),
        ("nthreads", ctypes.c_int),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_pt_solve.restype = ctypes.c_int
        L.oracle_pt_solve.argtypes = [ctypes.c_char_p, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_double)]
        L.oracle_pt_count_flops.restype = ctypes.c_int
        L.oracle_pt_count_flops.argtypes = [ctypes.c_char_p, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_longlong)]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a, typ):
    return a.ctypes.data_as(typ) if a is not None and a.size else ctypes.cast(None, typ)


def run(system, spec, nthreads=0, count_flops=False, _job_hook=None):
    """Integrate with the oracle.  `spec` keys: algorithm, nt, ndo, n_active,
    options (dict), rtol, atol, etol, edir, evcap, and the SoA arrays TD, X, T,
    cPAR, sPAR, sPARi, ACC, ACCi, DOidx, DOt, DOx (float64 / int32).  Returns a
    dict of output arrays (copies) plus 'seconds' (or 'flops')."""
    nt = int(spec["nt"])
    ndo = int(spec.get("ndo", 0))
    evcap = int(spec.get("evcap", 0))
    o = dict(dt0=1e-2, dtmax=1e6, dtmin=1e-12, grow=5.0, shrink=0.1, do_min_dt=0.0, do_freq=1)
    o.update(spec.get("options", {}))

    def f64(name, shape):
        a = spec.get(name)
        a = np.zeros(shape, np.float64) if a is None else np.array(a, dtype=np.float64, copy=True)
        return np.ascontiguousarray(a.reshape(shape))

    def i32(name, shape):
        a = spec.get(name)
        a = np.zeros(shape, np.int32) if a is None else np.array(a, dtype=np.int32, copy=True)
        return np.ascontiguousarray(a.reshape(shape))

    sd, ncp, na, nia, nsp, nisp, ne = (spec[k] for k in ("sd", "ncp", "na", "nia", "nsp", "nisp", "ne"))
    out = {
        "TD": f64("TD", (2, nt)), "X": f64("X", (sd, nt)), "T": f64("T", (nt,)), "cPAR": f64("cPAR", (ncp, nt)),
        "sPAR": f64("sPAR", (nsp,)), "sPARi": i32("sPARi", (nisp,)), "ACC": f64("ACC", (na, nt)),
        "ACCi": i32("ACCi", (nia, nt)), "DOidx": i32("DOidx", (nt,)), "DOt": f64("DOt", (ndo, nt)),
        "DOx": f64("DOx", (ndo, sd, nt)),
        "nacc": np.zeros(nt, np.int64), "nrej": np.zeros(nt, np.int64), "status": np.zeros(nt, np.int32),
        "nev": np.zeros(nt, np.int32), "evlog": np.zeros((evcap, sd + 1, nt), np.float64),
        "evlogidx": np.zeros((evcap, nt), np.int32),
    }
    rtol = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("rtol", 1e-8), np.float64), (sd,)))
    atol = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("atol", 1e-8), np.float64), (sd,)))
    etol = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("etol", 1e-6), np.float64), (max(ne, 1),)))
    edir = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("edir", 0), np.int32), (max(ne, 1),)))
    j = Job()
    j.nt, j.ndo, j.algorithm = nt, ndo, int(spec["algorithm"])
    j.n_active = int(spec.get("n_active", nt))
    for k in ("dt0", "dtmax", "dtmin", "grow", "shrink", "do_min_dt"):
        setattr(j, k, float(o[k]))
    j.do_freq = int(o["do_freq"])
    j.rtol, j.atol, j.etol, j.edir = _ptr(rtol, _dp), _ptr(atol, _dp), _ptr(etol, _dp), _ptr(edir, _ip)
    j.evcap = evcap
    for k in ("TD", "X", "T", "cPAR", "sPAR", "ACC", "DOt", "DOx", "evlog"):
        setattr(j, k, _ptr(out[k], _dp))
    for k in ("sPARi", "ACCi", "DOidx", "status", "nev", "evlogidx"):
        setattr(j, k, _ptr(out[k], _ip))
    j.nacc, j.nrej = _ptr(out["nacc"], _lp), _ptr(out["nrej"], _lp)
    j.nthreads = int(nthreads)
    if _job_hook is not None:
        _job_hook(j)
        return out
    if count_flops:
        fl = ctypes.c_longlong()
        rc = lib().oracle_pt_count_flops(system.encode(), ctypes.byref(j), ctypes.byref(fl))
        out["flops"] = fl.value
    else:
        sec = ctypes.c_double()
        rc = lib().oracle_pt_solve(system.encode(), ctypes.byref(j), ctypes.byref(sec))
        out["seconds"] = sec.value
    if rc != 0:
        raise RuntimeError(f"oracle failed for system {system!r} (rc={rc})")
    return out


def max_threads():
    return lib().oracle_max_threads()


def trace(system, spec, tid=0, cap=100000):
    """Per-attempt trace of one instance: arrays t (before the attempt), h,
    accepted flag and the controller's tol/err ratio."""
    L = lib()
    if not hasattr(L, "_trace_set"):
        L.oracle_pt_trace.restype = ctypes.c_int
        L.oracle_pt_trace.argtypes = [ctypes.c_char_p, ctypes.POINTER(Job), ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _ip, _dp, ctypes.POINTER(ctypes.c_int)]
        L._trace_set = True
    t = np.zeros(cap)
    h = np.zeros(cap)
    acc = np.zeros(cap, np.int32)
    ratio = np.zeros(cap)
    n = ctypes.c_int()
    # reuse run()'s marshalling through a one-instance spec
    holder = {}

    def hook(j):
        rc = L.oracle_pt_trace(system.encode(), ctypes.byref(j), int(tid), int(cap), _ptr(t, _dp), _ptr(h, _dp),
                               _ptr(acc, _ip), _ptr(ratio, _dp), ctypes.byref(n))
        if rc != 0:
            raise RuntimeError("oracle trace failed")
        holder["n"] = n.value

    run(system, spec, _job_hook=hook)
    k = holder["n"]
    return {"t": t[:k], "h": h[:k], "accepted": acc[:k], "ratio": ratio[:k]}


def sensitivity(system, spec, nthreads=0):
    """d x(T) / d (x(0), cPAR) on the frozen accepted-step sequence, shape
    (SD, SD + NCP, NT), plus the ordinary outputs of the same run."""
    L = lib()
    if not hasattr(L, "_sens_set"):
        L.oracle_pt_sensitivity.restype = ctypes.c_int
        L.oracle_pt_sensitivity.argtypes = [ctypes.c_char_p, ctypes.POINTER(Job), _dp]
        L._sens_set = True
    sd, ncp, nt = spec["sd"], spec["ncp"], spec["nt"]
    jac = np.zeros((sd, sd + ncp, nt))

    def hook(j):
        j.nthreads = int(nthreads)
        if L.oracle_pt_sensitivity(system.encode(), ctypes.byref(j), _ptr(jac, _dp)) != 0:
            raise RuntimeError("oracle sensitivity failed")

    out = run(system, spec, _job_hook=hook)
    out["jac"] = jac
    return out


# ---------------------------------------------------------------- per-block
class PBJob(ctypes.Structure):
    _fields_ = [
        ("ns", ctypes.c_int), ("ups", ctypes.c_int), ("n_active", ctypes.c_int), ("ndo", ctypes.c_int),
        ("algorithm", ctypes.c_int), ("cbw", ctypes.c_int), ("cci", ctypes.c_int), ("cmat_size", ctypes.c_int),
        ("dt0", ctypes.c_double), ("dtmax", ctypes.c_double), ("dtmin", ctypes.c_double),
        ("grow", ctypes.c_double), ("shrink", ctypes.c_double), ("do_min_dt", ctypes.c_double),
        ("do_freq", ctypes.c_int),
        ("rtol", _dp), ("atol", _dp), ("etol", _dp), ("edir", _ip),
        ("X", _dp), ("uPAR", _dp), ("uACC", _dp), ("uACCi", _ip),
        ("TD", _dp), ("T", _dp), ("sPAR", _dp), ("sACC", _dp), ("sACCi", _ip), ("strength", _dp), ("DOidx", _ip),
        ("DOt", _dp), ("DOx", _dp),
        ("gPAR", _dp), ("gPARi", _ip), ("cIndex", _ip), ("cMat", _dp),
        ("nacc", _lp), ("nrej", _lp), ("status", _ip),
        ("nthreads", ctypes.c_int),
    ]


def _pb_job(spec, out, nthreads=0):
    """PBJob over copies of the spec's arrays; `out` receives the arrays (the
    outputs are written there)."""
    ns, ups, ud = spec["ns"], spec["ups"], spec["ud"]
    nu = ns * ups
    ndo = spec.get("ndo", 0)
    o = dict(dt0=1e-2, dtmax=1e6, dtmin=1e-12, grow=5.0, shrink=0.1, do_min_dt=0.0, do_freq=1)
    o.update(spec.get("options", {}))

    def cp(name, dtype):
        return np.ascontiguousarray(np.array(spec[name], dtype=dtype, copy=True))

    out.update({k: cp(k, np.float64) for k in ("X", "uPAR", "uACC", "TD", "T", "sPAR", "sACC", "strength", "gPAR", "cMat")})
    out.update({k: cp(k, np.int32) for k in ("uACCi", "sACCi", "gPARi", "cIndex", "DOidx")})
    out["DOt"] = np.zeros((ndo, ns))
    out["DOx"] = np.zeros((ndo, ud, nu))
    out["nacc"] = np.zeros(ns, np.int64)
    out["nrej"] = np.zeros(ns, np.int64)
    out["status"] = np.zeros(ns, np.int32)
    ne = spec["ne"]
    out["_rtol"] = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("rtol", 1e-8), np.float64), (ud,)))
    out["_atol"] = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("atol", 1e-8), np.float64), (ud,)))
    out["_etol"] = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("etol", 1e-6), np.float64), (max(ne, 1),)))
    out["_edir"] = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.get("edir", 0), np.int32), (max(ne, 1),)))
    j = PBJob()
    j.ns, j.ups, j.n_active, j.ndo = ns, ups, int(spec.get("n_active", ns)), ndo
    j.algorithm, j.cbw, j.cci = int(spec["algorithm"]), int(spec["cbw"]), int(spec["cci"])
    j.cmat_size = out["cMat"].size // max(1, spec["nc"])
    for k in ("dt0", "dtmax", "dtmin", "grow", "shrink", "do_min_dt"):
        setattr(j, k, float(o[k]))
    j.do_freq = int(o["do_freq"])
    j.rtol, j.atol = _ptr(out["_rtol"], _dp), _ptr(out["_atol"], _dp)
    j.etol, j.edir = _ptr(out["_etol"], _dp), _ptr(out["_edir"], _ip)
    for k in ("X", "uPAR", "uACC", "TD", "T", "sPAR", "sACC", "strength", "DOt", "DOx", "gPAR", "cMat"):
        setattr(j, k, _ptr(out[k], _dp))
    for k in ("uACCi", "sACCi", "DOidx", "gPARi", "cIndex", "status"):
        setattr(j, k, _ptr(out[k], _ip))
    j.nacc, j.nrej = _ptr(out["nacc"], _lp), _ptr(out["nrej"], _lp)
    j.nthreads = int(nthreads)
    return j


def run_pb(system, spec, nthreads=0):
    """Integrate a per-block spec (see mpgos.workloads.new_pb_spec) with the
    oracle; returns the output arrays (copies) plus 'seconds'."""
    L = lib()
    if not hasattr(L, "_pb_set"):
        L.oracle_pb_solve.restype = ctypes.c_int
        L.oracle_pb_solve.argtypes = [ctypes.c_char_p, ctypes.POINTER(PBJob), ctypes.POINTER(ctypes.c_double)]
        L._pb_set = True
    out = {}
    j = _pb_job(spec, out, nthreads)
    sec = ctypes.c_double()
    if L.oracle_pb_solve(system.encode(), ctypes.byref(j), ctypes.byref(sec)) != 0:
        raise RuntimeError(f"oracle per-block failed for {system!r}")
    for k in ("_rtol", "_atol", "_etol", "_edir"):
        out.pop(k)
    out["seconds"] = sec.value
    return out


def sensitivity_pb(system, spec, nthreads=0):
    """Frozen-step forward sensitivities of a per-block spec (pb_sens.h):
    array [ns, N, UD, UPS] with directions N = (system parameters, initial
    state of unit 0, unit parameters of unit 0)."""
    L = lib()
    if not hasattr(L, "_pbsens_set"):
        L.oracle_pb_sensitivity.restype = ctypes.c_int
        L.oracle_pb_sensitivity.argtypes = [ctypes.c_char_p, ctypes.POINTER(PBJob), _dp, ctypes.POINTER(ctypes.c_int)]
        L._pbsens_set = True
    ns, ups, ud = spec["ns"], spec["ups"], spec["ud"]
    n = spec["nsp"] + ud + spec["nup"]
    jac = np.zeros((ns, n, ud, ups))
    keep = {}
    j = _pb_job(spec, keep, nthreads)
    nd = ctypes.c_int()
    if L.oracle_pb_sensitivity(system.encode(), ctypes.byref(j), _ptr(jac, _dp), ctypes.byref(nd)) != 0:
        raise RuntimeError(f"oracle per-block sensitivity failed for {system!r}")
    assert nd.value == n
    return jac
