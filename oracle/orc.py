"""ctypes loader of the CPU oracle liborc.so (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker of the HIP path.
"""
import ctypes as C
import os
import subprocess

from parmec_amd import pmx

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liborc.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liborc.so"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        lib = C.CDLL(LIBPATH)
        lib.orc_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        pmx._declare(lib, "orc_")
        lib.orc_expmap.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.orc_expmap.restype = None
        lib.orc_triangle_sphere.argtypes = [C.POINTER(C.c_double)] * 4 + [C.c_double] + [C.POINTER(C.c_double)] * 2
        lib.orc_triangle_sphere.restype = C.c_double
        lib.orc_tree_pairs.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                       C.c_longlong]
        lib.orc_tree_pairs.restype = C.c_longlong
        lib.orc_morton.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.orc_morton.restype = C.c_uint
        lib.orc_tree_leaves.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.orc_tree_leaves.restype = C.c_int
        lib.orc_tree_reach.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.c_longlong]
        lib.orc_tree_reach.restype = C.c_longlong
        lib.orc_trig.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        lib.orc_trig.restype = None
        lib.orc_time_step.argtypes = [C.POINTER(C.c_double), C.c_double, C.c_double, C.POINTER(C.c_double),
                                      C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        lib.orc_time_step.restype = C.c_double
        lib.orc_extremum_eigenvalue.argtypes = [C.POINTER(C.c_double), C.c_int]
        lib.orc_extremum_eigenvalue.restype = C.c_double
        _lib = lib
    return _lib


class Oracle(pmx.Engine):
    """Same interface as pmx.Engine, backed by the CPU restatement."""

    prefix = "orc_"

    def __init__(self, nthreads=0):
        self.lib = load()
        h = C.c_void_p()
        self._call("create", nthreads, C.byref(h))
        self.h = h


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def trig(fn, x):
    """pmx_trig.h's sin (fn 0), cos (1) or acos (2) of an array."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    load().orc_trig(int(fn), _dp(x), _dp(y), int(x.size))
    return y


def time_step(kact, kmax, emax, krot, mass, inertia):
    """time_step_task for one particle: (h, omega, ratio)."""
    import numpy as np
    ka = np.ascontiguousarray(kact, dtype=np.float64)
    kr = np.ascontiguousarray(krot, dtype=np.float64)
    J = np.ascontiguousarray(inertia, dtype=np.float64).reshape(9)
    o, r = C.c_double(0.0), C.c_double(0.0)
    h = load().orc_time_step(_dp(ka), float(kmax), float(emax), _dp(kr), float(mass), _dp(J), C.byref(o), C.byref(r))
    return h, o.value, r.value


def extremum_eigenvalue(A, minimum):
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    return load().orc_extremum_eigenvalue(_dp(A), int(minimum))


def skyline_lu(A, b):
    """amgcl's skyline_lu solve of a dense matrix (None on a zero pivot)."""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    lib = load()
    lib.orc_skyline_lu.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return None if lib.orc_skyline_lu(len(b), _dp(A), _dp(b), _dp(x)) else x


def cuthill_mckee(ptr, col):
    """amgcl's Cuthill-McKee order of a CSR pattern."""
    import numpy as np
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    n = len(ptr) - 1
    perm = np.zeros(n, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    load().orc_cuthill_mckee(n, ptr.ctypes.data_as(ip), col.ctypes.data_as(ip), perm.ctypes.data_as(ip))
    return perm
