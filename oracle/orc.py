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
