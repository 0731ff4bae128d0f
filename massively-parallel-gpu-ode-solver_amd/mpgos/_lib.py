"""ctypes binding of libmpgos.so (the C ABI in include/mpgos.h).

The library is the product: it holds the gfx950 kernels.  Loading fails loudly
when it has not been built; there is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmpgos.so")

_lib = None


class MPGOSError(RuntimeError):
    """A libmpgos call failed (the reference would have called exit())."""


class _Dims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nt", "sd", "ncp", "nsp", "nisp", "ne", "na", "nia", "ndo", "algorithm", "precision", "has_vjp")]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MPGOSError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                         "(or make -C massively-parallel-gpu-ode-solver_amd)")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_double, c_void_p, c_char_p, c_size_t, c_long = (
        ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_long)
    H = c_void_p
    sig = {
        "mpgos_last_error": (c_char_p, []),
        "mpgos_pt_list_systems": (c_char_p, []),
        "mpgos_pt_create": (c_int, [c_char_p, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(H)]),
        "mpgos_pt_destroy": (c_int, [H]),
        "mpgos_pt_get_dims": (c_int, [H, ctypes.POINTER(_Dims)]),
        "mpgos_pt_solver_option": (c_int, [H, c_int, c_int, c_double]),
        "mpgos_pt_set_host": (c_int, [H, c_int, c_int, c_int, c_int, c_double]),
        "mpgos_pt_get_host": (c_int, [H, c_int, c_int, c_int, c_int, ctypes.POINTER(c_double)]),
        "mpgos_pt_host_buffer": (c_void_p, [H, c_int, ctypes.POINTER(c_size_t)]),
        "mpgos_pt_device_buffer": (c_void_p, [H, c_int, ctypes.POINTER(c_size_t)]),
        "mpgos_pt_sync_h2d": (c_int, [H, c_int]),
        "mpgos_pt_sync_d2h": (c_int, [H, c_int]),
        "mpgos_pt_solve": (c_int, [H]),
        "mpgos_pt_solve_slice": (c_int, [H, ctypes.c_longlong, ctypes.POINTER(c_int)]),
        "mpgos_pt_solve_noalloc": (c_int, [H]),
        "mpgos_pt_synchronise_device": (c_int, [H]),
        "mpgos_pt_insert_synchronisation_point": (c_int, [H]),
        "mpgos_pt_synchronise_solver": (c_int, [H]),
        "mpgos_pt_print": (c_int, [H, c_int, c_int, c_char_p]),
        "mpgos_pt_last_solve_ms": (c_int, [H, ctypes.POINTER(c_double)]),
        "mpgos_pt_stream": (c_void_p, [H]),
        "mpgos_pt_adjoint_record": (c_int, [H, c_int, c_int]),
        "mpgos_pt_adjoint_reverse": (c_int, [H]),
        "mpgos_math_eval_host": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_long]),
        "mpgos_math_eval_device": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_long]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def hip_runtime_path():
    """Path of the HIP runtime libmpgos.so is linked against (after lib() has
    loaded it).  Other packages may bring their own copy of libamdhip64 into
    the process; calls on buffers of this library must go to this one."""
    lib()
    candidates = []
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 6 and "libamdhip64" in parts[-1]:
                    candidates.append(parts[-1])
    except OSError:
        pass
    for p in candidates:
        if p.startswith("/opt/rocm"):
            return p
    return candidates[0] if candidates else "libamdhip64.so.7"


def check(rc):
    if rc != 0:
        msg = lib().mpgos_last_error().decode()
        raise MPGOSError(f"libmpgos error {rc}: {msg}")
    return rc
