"""ctypes binding of libpmx.so (include/pmx.h) and the SoA model container.

`Model` owns numpy arrays laid out like Parmec's global buffers
(parmec.h:66-283): a k-component per-item array has shape (k, n), C order,
so component c of item i sits at c*n + i.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpmx.so")

PMX_OK = 0
PMX_FP64 = 0
PMX_MIXED = 1
NIPARAM = 7
SPRING, DAMPER, FRISTAT, FRIDYN, FRIROL, FRIDRIL, KSKN = range(7)
GRANULAR_FORCE = 0
SPRING_NONLINEAR_ELASTIC, SPRING_GENERAL_NONLINEAR = 0, 1
SPRING_YIELDED, SPRING_UNLOADING, SPRDIR_CONSTANT, SPRDIR_PLANAR, SPRDIR_FOLLOWER, SPRDIR_PROJECT = 1, 2, 4, 8, 16, 32

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class pmx_model(C.Structure):
    _fields_ = [
        ("parnum", C.c_int), ("angular", _dp), ("linear", _dp), ("rotation", _dp), ("position", _dp),
        ("inertia", _dp), ("inverse", _dp), ("mass", _dp), ("invm", _dp), ("force", _dp), ("torque", _dp),
        ("flags", _ip),
        ("ellnum", C.c_int), ("center", _dp), ("radii", _dp), ("part", _ip), ("ellcol", _ip),
        ("trinum", C.c_int), ("tri", _dp), ("tricol", _ip), ("triobs", _ip),
        ("obsnum", C.c_int), ("trirng", _ip), ("obspnt", _dp), ("obslin", _dp), ("obsang", _dp),
        ("nodnum", C.c_int), ("nodes", _dp), ("nodpart", _ip), ("facnum", C.c_int), ("facnod", _ip), ("factri", _ip),
        ("pairnum", C.c_int), ("pairs", _ip), ("ikind", _ip), ("iparam", _dp),
        ("sprnum", C.c_int), ("sprtype", _ip), ("sprpart", _ip), ("sprpnt", _dp), ("spridx", _ip), ("spring", _dp),
        ("sprcrv", C.c_int), ("dashidx", _ip), ("dashpot", _dp), ("dashcrv", C.c_int), ("unidx", _ip),
        ("unload", _dp), ("uncrv", C.c_int), ("yield_", _dp), ("sprdir", _dp), ("sprflg", _ip), ("sprfric", _dp),
        ("sprkskn", _dp), ("sprsdsp", _dp), ("stroke0", _dp), ("stroke", _dp), ("sprfrc", _dp), ("unspring", _ip),
        ("sproffset", _ip),
        ("prsnum", C.c_int), ("prspart", _ip), ("prskind", _ip), ("prscol", _ip),
        ("curtime", C.c_double), ("curstep", C.c_double), ("stepnum", C.c_int),
        ("lcnum", C.c_int), ("lcidx", _ip), ("lcurve", _dp), ("lccrv", C.c_int),
        ("trqsprnum", C.c_int), ("trqsprpart", _ip), ("trqzdir0", _dp), ("trqxdir0", _dp), ("krpyidx", _ip),
        ("krpy", _dp), ("krpycrv", C.c_int), ("drpyidx", _ip), ("drpy", _dp), ("drpycrv", C.c_int), ("trqcone", _ip),
        ("trqzdir1", _dp), ("trqxdir1", _dp), ("trqrpy", _dp), ("trqrpytot", _dp), ("trqrpyspr", _dp),
        ("unsprnum", C.c_int), ("tspridx", _ip), ("tsprings", _ip), ("mspridx", _ip), ("msprings", _ip),
        ("actidx", _ip), ("activate", _ip), ("unlim", _dp), ("unent", _ip), ("unop", _ip), ("unabs", _ip),
        ("nsteps", _ip), ("nfreq", _ip), ("unaction", _ip),
        ("rstnum", C.c_int), ("rstpart", _ip), ("rstlin", _dp), ("rstang", _dp),
        ("jnum", C.c_int), ("jpart", _ip), ("jpoint", _dp), ("jreac", _dp),
    ]

# UNSPRING operators (constants.h:70) and torsion spring cone flags (66-68)
OP_SUM, OP_MAX, OP_MIN = 0, 1, 2
TRQCONE_ROLL, TRQCONE_PITCH, TRQCONE_YAW = 1, 2, 4


class pmx_step_inputs(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("damping", C.c_double * 6), ("obsvel", _dp), ("prs", _dp),
                ("hisrow", C.c_int), ("output", C.c_int)]


class pmx_forcing(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("gtab", _dp), ("damping", C.c_double * 6), ("dtab", _dp),
                ("obstab", _dp), ("obstab0", _dp), ("prstab", _dp), ("prscols", C.c_int),
                ("hisrow", _ip), ("hisrows", C.c_int), ("hisout", _dp),
                ("outstep", _ip), ("noutstep", C.c_int), ("out_fn", C.c_void_p), ("out_user", C.c_void_p),
                ("step_fn", C.c_void_p), ("step_user", C.c_void_p)]


# pmx_forcing.out_fn: void (*)(void *user, int step)
OUT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int)
# pmx_forcing.step_fn: int (*)(void *user, int k, double curtime, double step0, pmx_step_inputs *in)
STEP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.POINTER(pmx_step_inputs))

# pmx_critical estimates
CRIT_PARTICLE, CRIT_SPRING = 0, 1


class StepInputs:
    """The host inputs of one adaptive step, as on_step(k, curtime, step0,
    inputs) fills them: gravity (3), damping (6), obsvel [obsnum][6]
    (angular, linear velocity at curtime + step0), prs [prscols][6]
    (prescribed linear, angular values at curtime), hisrow (the history row
    sampling the step's end state, or -1) and output (1: output point)."""

    def __init__(self, raw, obsnum, prscols):
        self._raw = raw
        self.gravity = np.ctypeslib.as_array(raw.gravity)
        self.damping = np.ctypeslib.as_array(raw.damping)
        self.obsvel = np.ctypeslib.as_array(raw.obsvel, shape=(max(obsnum, 1) * 6,)).reshape(-1, 6)[:obsnum]
        self.prs = np.ctypeslib.as_array(raw.prs, shape=(max(prscols, 1) * 6,)).reshape(-1, 6)[:prscols]
        self.hisrow = -1
        self.output = 0


# time-history entities of HISTORY (input.cpp:3177-3281), as PMX_HIS_* in pmx.h
HIS_ENTITIES = ["TIME", "PX", "PY", "PZ", "|P|", "DX", "DY", "DZ", "|D|", "VX", "VY", "VZ", "|V|", "OX", "OY", "OZ",
                "|O|", "FX", "FY", "FZ", "|F|", "TX", "TY", "TZ", "|T|", "LENGTH", "STROKE", "F", "SF", "FF", "SS",
                "ZDIR_X", "ZDIR_Y", "ZDIR_Z", "XDIR_X", "XDIR_Y", "XDIR_Z", "ROLL", "PITCH", "YAW", "TRQTOT_R",
                "TRQTOT_P", "TRQTOT_Y", "TRQSPR_R", "TRQSPR_P", "TRQSPR_Y", "JREAC_X", "JREAC_Y", "JREAC_Z", "|JREAC|"]
HIS_SPRING_FIRST = HIS_ENTITIES.index("LENGTH")
HIS_TORSION_FIRST = HIS_ENTITIES.index("ZDIR_X")
HIS_JOINT_FIRST = HIS_ENTITIES.index("JREAC_X")


class pmx_history(C.Structure):
    _fields_ = [("count", C.c_int), ("entity", _ip), ("off", _ip), ("list", _ip), ("haspoint", _ip), ("point", _dp)]


class pmx_options(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("row_capacity", C.c_int), ("verbose", C.c_int),
                ("tune", C.c_int), ("det_waves", C.c_int), ("rows_block", C.c_int), ("critical", C.c_int)]


class pmx_pair(C.Structure):
    _fields_ = [("master", C.c_int), ("slave", C.c_int), ("state", C.c_double * 3)]


class pmx_stats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("rebuilds", C.c_longlong), ("pairs", C.c_longlong), ("leaves", C.c_int),
                ("max_row", C.c_int), ("seconds", C.c_double), ("det_waves", C.c_int), ("rows_block", C.c_int),
                ("leaf_slots", C.c_int), ("partial_rebuilds", C.c_longlong),
                ("graph_timed", C.c_int), ("grown", C.c_longlong), ("row_capacity", C.c_int),
                ("busy", C.c_longlong), ("window_ms", C.c_double)]


class pmx_kernel_time(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_longlong), ("ms", C.c_double), ("bytes", C.c_double)]


EXPORTS = ["pmx_step_times", "pmx_create", "pmx_upload", "pmx_dem", "pmx_download", "pmx_contacts",
           "pmx_get_stats", "pmx_destroy", "pmx_last_error", "pmx_set_profiling", "pmx_get_profile",
           "pmx_set_history", "pmx_history_sample"]


class pmx_server_options(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("row_capacity", C.c_int), ("max_batch", C.c_int),
                ("max_active", C.c_int), ("gather_s", C.c_double), ("hbm_budget", C.c_size_t),
                ("repack_s", C.c_double), ("longest_first", C.c_int)]


class pmx_request_info(C.Structure):
    _fields_ = [("id", C.c_longlong), ("status", C.c_int), ("batch", C.c_int), ("steps", C.c_longlong),
                ("particles", C.c_longlong), ("pairs", C.c_longlong), ("bytes", C.c_size_t), ("submit_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double)]


class pmx_server_stats(C.Structure):
    _fields_ = [("queued", C.c_longlong), ("running", C.c_longlong), ("done", C.c_longlong),
                ("failed", C.c_longlong), ("batches", C.c_longlong), ("active_batches", C.c_int),
                ("bytes_in_use", C.c_size_t), ("particle_steps", C.c_double), ("seconds", C.c_double)]


PMX_EBUSY = -7
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class PmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pmx error {code}: {msg}")
        self.code = code


def _declare(lib, prefix):
    f = lambda n: getattr(lib, prefix + n)
    f("step_times").argtypes = [C.c_double, C.c_double, C.c_double, _dp, C.c_int, _ip]
    f("last_error").restype = C.c_char_p
    f("upload").argtypes = [C.c_void_p, C.POINTER(pmx_model)]
    f("download").argtypes = [C.c_void_p, C.POINTER(pmx_model)]
    f("dem").argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(pmx_forcing), _dp]
    f("contacts").argtypes = [C.c_void_p, C.POINTER(pmx_pair), C.c_size_t, C.POINTER(C.c_size_t)]
    f("get_stats").argtypes = [C.c_void_p, C.POINTER(pmx_stats)]
    f("destroy").argtypes = [C.c_void_p]
    f("destroy").restype = None
    f("critical").argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
    f("dem_adaptive").argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(pmx_forcing), _dp]
    return lib


_lib = None


def load(path=LIBPATH):
    """Load libpmx.so; raises when the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python parmec_amd/build.py`")
        lib = C.CDLL(path)
        lib.pmx_create.argtypes = [C.POINTER(pmx_options), C.POINTER(C.c_void_p)]
        lib.pmx_upload_batch.argtypes = [C.c_void_p, C.POINTER(pmx_model), C.c_int]
        lib.pmx_download_batch.argtypes = [C.c_void_p, C.POINTER(pmx_model), C.c_int]
        lib.pmx_dem_batch.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(pmx_forcing), _dp]
        lib.pmx_contacts_instance.argtypes = [C.c_void_p, C.c_int, C.POINTER(pmx_pair), C.c_size_t,
                                              C.POINTER(C.c_size_t)]
        lib.pmx_download_instance.argtypes = [C.c_void_p, C.c_int, C.POINTER(pmx_model)]
        lib.pmx_server_create.argtypes = [C.POINTER(pmx_server_options), C.POINTER(C.c_void_p)]
        lib.pmx_server_submit.argtypes = [C.c_void_p, C.POINTER(pmx_model), C.c_double, C.c_double,
                                          C.POINTER(pmx_forcing), C.POINTER(C.c_longlong)]
        lib.pmx_server_wait.argtypes = [C.c_void_p, C.c_longlong, C.c_double, C.POINTER(pmx_model),
                                        C.POINTER(pmx_request_info)]
        lib.pmx_server_get_stats.argtypes = [C.c_void_p, C.POINTER(pmx_server_stats)]
        lib.pmx_server_destroy.argtypes = [C.c_void_p]
        lib.pmx_server_destroy.restype = None
        _lib = _declare(lib, "pmx_")
    return _lib


def step_times(duration, step, curtime=0.0, lib=None, prefix="pmx_"):
    """Curtime at the start of each step of dem(duration, step) (parmec.cpp:2418)."""
    lib = lib or load()
    f = getattr(lib, prefix + "step_times")
    n = C.c_int(0)
    f(duration, step, curtime, None, 0, C.byref(n))
    times = np.zeros(max(n.value, 1))
    f(duration, step, curtime, times.ctypes.data_as(_dp), n.value, C.byref(n))
    return times[: n.value]


def _ptr(a, kind):
    if a is None:
        return None
    return a.ctypes.data_as(_dp if kind == "d" else _ip)


class Model:
    """SoA state of a Parmec model (the arrays the step path reads and writes)."""

    def __init__(self, parnum=0, ellnum=0, trinum=0, pairnum=1):
        self.parnum, self.ellnum, self.trinum, self.pairnum = parnum, ellnum, trinum, pairnum
        N, E, T, P = parnum, ellnum, trinum, pairnum
        z = lambda *s: np.zeros(s, dtype=np.float64)
        zi = lambda *s: np.zeros(s, dtype=np.int32)
        self.angular, self.linear, self.rotation, self.position = z(6, N), z(3, N), z(9, N), z(6, N)
        self.rotation[[0, 4, 8]] = 1.0
        self.inertia, self.inverse, self.mass, self.invm = z(9, N), z(9, N), z(N), z(N)
        self.force, self.torque, self.flags = z(3, N), z(3, N), zi(N)
        self.center, self.radii, self.part, self.ellcol = z(6, E), z(3, E), zi(E), zi(E)
        self.tri, self.tricol, self.triobs = z(9, T), zi(T), np.full(T, -1, dtype=np.int32)
        self.pairs, self.ikind, self.iparam = zi(2 * P), zi(P), z(NIPARAM, P)
        self.iparam[DAMPER, 0] = 1.0
        self.iparam[KSKN, 0] = 0.5
        self.curtime, self.curstep, self.stepnum = 0.0, 0.0, 0
        self.set_mesh(0, 0)
        self.set_springs(0, 0, 0, 0)
        self.set_obstacles(0)
        self.set_prescribed(0)
        self.set_lcurves([])
        self.set_torsion(0, 0, 0)
        self.set_unsprings(0, 0, 0, 0)
        self.set_restraints(0)
        self.set_joints(0)

    def set_lcurves(self, curves):
        """Load curves (lcurve_from_time_series, parmec.cpp:1239-1269): a list
        of (x list, y list); lcidx [lcnum+1], lcurve [2][lccrv]."""
        self.lcnum = len(curves)
        self.lccrv = sum(len(x) for x, _ in curves)
        self.lcidx = np.zeros(self.lcnum + 1, dtype=np.int32)
        self.lcurve = np.zeros((2, self.lccrv))
        k = 0
        for i, (x, y) in enumerate(curves):
            self.lcidx[i] = k
            self.lcurve[0, k:k + len(x)], self.lcurve[1, k:k + len(x)] = x, y
            k += len(x)
        self.lcidx[self.lcnum] = k

    def set_torsion(self, trqsprnum, krpycrv, drpycrv):
        """Torsion springs (parmec.cpp:180-198): parts [2][n], reference axes
        [3][n], curve index ranges [3][n+1] into the shared krpy, drpy tables
        [2][crv], cone flags, and the state and outputs [3][n]."""
        T = trqsprnum
        self.trqsprnum, self.krpycrv, self.drpycrv = T, krpycrv, drpycrv
        self.trqsprpart = np.zeros((2, T), dtype=np.int32)
        self.trqzdir0, self.trqxdir0 = np.zeros((3, T)), np.zeros((3, T))
        self.krpyidx, self.drpyidx = np.zeros((3, T + 1), dtype=np.int32), np.zeros((3, T + 1), dtype=np.int32)
        self.krpy, self.drpy = np.zeros((2, krpycrv)), np.zeros((2, drpycrv))
        self.trqcone = np.zeros(T, dtype=np.int32)
        self.trqzdir1, self.trqxdir1, self.trqrpy = np.zeros((3, T)), np.zeros((3, T)), np.zeros((3, T))
        self.trqrpytot, self.trqrpyspr = np.zeros((3, T)), np.zeros((3, T))

    def set_unsprings(self, unsprnum, ntest, nswitch, nact):
        """UNSPRING conditions (input.cpp:2083-2295): tested, switched-off and
        activated spring lists (index ranges [n+1]), limits [2][n], entity,
        operator, abs, nsteps, nfreq (state) and action per condition."""
        U = unsprnum
        self.unsprnum = U
        self.tspridx, self.tsprings = np.zeros(U + 1, dtype=np.int32), np.zeros(ntest, dtype=np.int32)
        self.mspridx, self.msprings = np.zeros(U + 1, dtype=np.int32), np.zeros(nswitch, dtype=np.int32)
        self.actidx, self.activate = np.zeros(U + 1, dtype=np.int32), np.zeros(nact, dtype=np.int32)
        self.unlim = np.zeros((2, U))
        self.unent, self.unop, self.unabs = (np.zeros(U, dtype=np.int32) for _ in range(3))
        self.nsteps, self.nfreq = np.ones(U, dtype=np.int32), np.ones(U, dtype=np.int32)
        self.unaction = np.full(U, -1, dtype=np.int32)

    def set_restraints(self, rstnum):
        """RESTRAIN entries: particle, linear and angular directions [9][n]."""
        self.rstnum = rstnum
        self.rstpart = np.zeros(rstnum, dtype=np.int32)
        self.rstlin, self.rstang = np.zeros((9, rstnum)), np.zeros((9, rstnum))

    def set_joints(self, jnum):
        """BALL_JOINT entries: parts [2][n] (part2 -1: ground), point [3][n],
        reactions [3][n] (output)."""
        self.jnum = jnum
        self.jpart = np.full((2, jnum), -1, dtype=np.int32)
        self.jpoint, self.jreac = np.zeros((3, jnum)), np.zeros((3, jnum))

    def set_prescribed(self, prsnum):
        """Prescribed motion entries (parmec.h:244-251): particle, and the kind
        of the linear and angular parts [2][prsnum] (-1 none, 0 velocity,
        1 acceleration)."""
        self.prsnum = prsnum
        self.prspart = np.zeros(prsnum, dtype=np.int32)
        self.prskind = np.full((2, prsnum), -1, dtype=np.int32)
        self.prscol = np.arange(prsnum, dtype=np.int32)

    def set_obstacles(self, obsnum):
        """Moving obstacles (parmec.h:104-112): triangle ranges [obsnum][2],
        point, linear and angular velocity [obsnum][3]."""
        self.obsnum = obsnum
        self.trirng = np.zeros((obsnum, 2), dtype=np.int32)
        self.obspnt = np.zeros((obsnum, 3))
        self.obslin = np.zeros((obsnum, 3))
        self.obsang = np.zeros((obsnum, 3))

    def set_mesh(self, nodnum, facnum):
        """Mesh nodes [6][nodnum] (current, reference), node -> particle, and
        faces [3][facnum] of node numbers with face -> triangle."""
        self.nodnum, self.facnum = nodnum, facnum
        self.nodes = np.zeros((6, nodnum))
        self.nodpart = np.zeros(nodnum, dtype=np.int32)
        self.facnod = np.zeros((3, facnum), dtype=np.int32)
        self.factri = np.zeros(facnum, dtype=np.int32)

    def set_springs(self, sprnum, sprcrv, dashcrv, uncrv):
        """SPRING arrays of parmec.h:152-200; curve tables are [2][n] (x, y)."""
        S = sprnum
        self.sprnum, self.sprcrv, self.dashcrv, self.uncrv = S, sprcrv, dashcrv, uncrv
        self.sprtype = np.zeros(S, dtype=np.int32)
        self.sprpart = np.zeros((2, S), dtype=np.int32)
        self.sprpnt = np.zeros((2, 6, S))
        self.spridx = np.zeros(S + 1, dtype=np.int32)
        self.spring = np.zeros((2, sprcrv))
        self.dashidx = np.zeros(S + 1, dtype=np.int32)
        self.dashpot = np.zeros((2, dashcrv))
        self.unidx = np.zeros(S + 1, dtype=np.int32)
        self.unload = np.zeros((2, uncrv))
        self.yield_ = np.zeros((2, S))
        self.sprdir = np.zeros((6, S))
        self.sprflg = np.zeros(S, dtype=np.int32)
        self.sprfric = np.zeros(S)
        self.sprkskn = np.zeros(S)
        self.sprsdsp = np.zeros((3, S))
        self.stroke0 = np.zeros(S)
        self.stroke = np.zeros((3, S))
        self.sprfrc = np.zeros((3, S))
        self.unspring = np.full(S, -3, dtype=np.int32)
        self.sproffset = np.full(S, -1, dtype=np.int32)

    def copy(self):
        m = Model.__new__(Model)
        for k, v in self.__dict__.items():
            setattr(m, k, v.copy() if isinstance(v, np.ndarray) else v)
        return m

    _DBL = ("angular", "linear", "rotation", "position", "inertia", "inverse", "mass", "invm", "force", "torque",
            "center", "radii", "tri", "iparam", "nodes", "sprpnt", "spring", "dashpot", "unload", "yield_", "sprdir",
            "sprfric", "sprkskn", "sprsdsp", "stroke0", "stroke", "sprfrc", "obspnt", "obslin", "obsang",
            "lcurve", "trqzdir0", "trqxdir0", "krpy", "drpy", "trqzdir1", "trqxdir1", "trqrpy", "trqrpytot", "trqrpyspr",
            "unlim", "rstlin", "rstang", "jpoint", "jreac")
    _INT = ("flags", "part", "ellcol", "tricol", "triobs", "pairs", "ikind", "nodpart", "facnod", "factri", "sprtype",
            "sprpart", "spridx", "dashidx", "unidx", "sprflg", "unspring", "sproffset", "trirng", "prspart", "prskind", "prscol",
            "lcidx", "trqsprpart", "krpyidx", "drpyidx", "trqcone", "tspridx", "tsprings", "mspridx", "msprings",
            "actidx", "activate", "unent", "unop", "unabs", "nsteps", "nfreq", "unaction", "rstpart", "jpart")

    def check(self):
        for name in self._DBL:
            a = getattr(self, name)
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                setattr(self, name, np.ascontiguousarray(a, dtype=np.float64))
        for name in self._INT:
            a = getattr(self, name)
            if a.dtype != np.int32 or not a.flags.c_contiguous:
                setattr(self, name, np.ascontiguousarray(a, dtype=np.int32))

    def struct(self):
        self.check()
        s = pmx_model()
        s.parnum, s.ellnum, s.trinum, s.pairnum = self.parnum, self.ellnum, self.trinum, self.pairnum
        s.nodnum, s.facnum, s.sprnum = self.nodnum, self.facnum, self.sprnum
        s.obsnum = getattr(self, "obsnum", 0)
        s.prsnum = getattr(self, "prsnum", 0)
        s.sprcrv, s.dashcrv, s.uncrv = self.sprcrv, self.dashcrv, self.uncrv
        s.lcnum, s.lccrv = self.lcnum, self.lccrv
        s.trqsprnum, s.krpycrv, s.drpycrv = self.trqsprnum, self.krpycrv, self.drpycrv
        s.unsprnum, s.rstnum, s.jnum = self.unsprnum, self.rstnum, self.jnum
        for name in self._DBL:
            setattr(s, name, _ptr(getattr(self, name), "d"))
        for name in self._INT:
            setattr(s, name, _ptr(getattr(self, name), "i"))
        s.curtime, s.curstep, s.stepnum = self.curtime, self.curstep, self.stepnum
        self._keep = s
        return s

    def take_time(self, s):
        self.curtime, self.curstep, self.stepnum = s.curtime, s.curstep, s.stepnum


def forcing(nsteps=0, gravity=(0.0, 0.0, 0.0), gtab=None, damping=(0.0,) * 6, dtab=None, obstab=None, obstab0=None,
            prstab=None, hisrow=None, hisout=None, outsteps=None, on_output=None, on_step=None, obsnum=0,
            prscols=0):
    """Per-step inputs of dem(): gravity and damping constants or tables
    ([nsteps][3], [nsteps][6]); moving obstacle velocities (angular 0-2,
    linear 3-5) per step at curtime+step ([nsteps][obsnum][6]) and at step/2
    for the first call ([obsnum][6]).  For dem_adaptive, on_step(k, curtime,
    step0, inputs) fills each step's inputs (StepInputs; obsnum and prscols
    size its arrays); an exception there stops the call."""
    f = pmx_forcing()
    for k in range(3):
        f.gravity[k] = gravity[k]
    for k in range(6):
        f.damping[k] = damping[k]
    keep = []
    if gtab is not None:
        gtab = np.ascontiguousarray(gtab, dtype=np.float64).reshape(-1, 3)
        assert gtab.shape[0] >= nsteps
        f.gtab = gtab.ctypes.data_as(_dp)
        keep.append(gtab)
    if dtab is not None:
        dtab = np.ascontiguousarray(dtab, dtype=np.float64).reshape(-1, 6)
        assert dtab.shape[0] >= nsteps
        f.dtab = dtab.ctypes.data_as(_dp)
        keep.append(dtab)
    if obstab is not None:
        obstab = np.ascontiguousarray(obstab, dtype=np.float64)
        assert obstab.shape[0] >= nsteps and obstab.shape[-1] == 6
        f.obstab = obstab.ctypes.data_as(_dp)
        keep.append(obstab)
    if obstab0 is not None:
        obstab0 = np.ascontiguousarray(obstab0, dtype=np.float64)
        f.obstab0 = obstab0.ctypes.data_as(_dp)
        keep.append(obstab0)
    if prstab is not None:
        prstab = np.ascontiguousarray(prstab, dtype=np.float64)
        assert prstab.ndim == 3 and prstab.shape[0] >= nsteps and prstab.shape[-1] == 6
        f.prstab = prstab.ctypes.data_as(_dp)
        f.prscols = prstab.shape[1]
        keep.append(prstab)
    if hisrow is not None:
        # the history row sampling each step's end state (-1: none) and the
        # output array [rows][histories] that dem() fills
        hisrow = np.ascontiguousarray(hisrow, dtype=np.int32)
        assert hisrow.shape[0] >= nsteps and hisout is not None and hisout.dtype == np.float64
        assert hisout.flags["C_CONTIGUOUS"]
        f.hisrow = hisrow.ctypes.data_as(_ip)
        f.hisrows = hisout.shape[0]
        f.hisout = hisout.ctypes.data_as(_dp)
        keep.extend([hisrow, hisout])
    if outsteps is not None and len(outsteps) and on_output is not None:
        # pause after these steps of the call (0-based) and call on_output(step)
        outsteps = np.ascontiguousarray(outsteps, dtype=np.int32)
        cb = OUT_FN(lambda user, step: on_output(int(step)))
        f.outstep = outsteps.ctypes.data_as(_ip)
        f.noutstep = len(outsteps)
        f.out_fn = C.cast(cb, C.c_void_p)
        keep.extend([outsteps, cb])
    if on_step is not None:
        errors = []

        def step_cb(user, k, curtime, step0, raw):
            try:
                inp = StepInputs(raw.contents, obsnum, prscols)
                on_step(int(k), float(curtime), float(step0), inp)
                raw.contents.hisrow = int(inp.hisrow)
                raw.contents.output = int(inp.output)
                return 0
            except BaseException as e:   # reported after the call returns
                errors.append(e)
                return 1
        cb = STEP_FN(step_cb)
        f.step_fn = C.cast(cb, C.c_void_p)
        f._step_errors = errors
        keep.append(cb)
        if on_output is not None and not (outsteps is not None and len(outsteps)):
            ocb = OUT_FN(lambda user, step: on_output(int(step)))
            f.out_fn = C.cast(ocb, C.c_void_p)
            keep.append(ocb)
    f._keep = keep
    return f


class Engine:
    """One device context of libpmx (pmx_create ... pmx_destroy)."""

    prefix = "pmx_"

    def __init__(self, precision=PMX_FP64, device=0, row_capacity=0, lib=None, tune=0, det_waves=0, rows_block=0,
                 critical=0):
        self.lib = lib or load()
        opt = pmx_options(precision, device, row_capacity, 0, tune, det_waves, rows_block, critical)
        h = C.c_void_p()
        self._call("create", C.byref(opt), C.byref(h))
        self.h = h

    def _call(self, name, *args):
        rc = getattr(self.lib, self.prefix + name)(*args)
        if rc != PMX_OK:
            raise PmxError(rc, getattr(self.lib, self.prefix + "last_error")().decode())
        return rc

    def set_history(self, entities, lists, points=None):
        """Time histories of the uploaded model (pmx_set_history): entity
        names of HIS_ENTITIES, a list of particle (spring) numbers each, and
        per history a body point (3-tuple) or None."""
        n = len(entities)
        ent = np.array([HIS_ENTITIES.index(e) for e in entities], dtype=np.int32)
        off = np.zeros(n + 1, dtype=np.int32)
        for i, l in enumerate(lists):
            off[i + 1] = off[i] + len(l)
        lst = np.array([k for l in lists for k in l] or [0], dtype=np.int32)
        hp = np.zeros(max(n, 1), dtype=np.int32)
        pt = np.zeros((max(n, 1), 3))
        for i, p in enumerate(points or [None] * n):
            if p is not None:
                hp[i] = 1
                pt[i] = [float(x) for x in p]
        h = pmx_history(n, ent.ctypes.data_as(_ip), off.ctypes.data_as(_ip), lst.ctypes.data_as(_ip),
                        hp.ctypes.data_as(_ip), pt.ctypes.data_as(_dp))
        self._call("set_history", self.h, C.byref(h))
        self._nhis = n

    def energy(self, instance=-1, gravity=(0.0, 0.0, 0.0)):
        """pmx_energy: (kinetic translational, kinetic rotational, potential,
        momentum x y z, angular momentum x y z) of the resident state."""
        out = np.zeros(9)
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        f = getattr(self.lib, self.prefix + "energy")
        f.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        self._call("energy", self.h, int(instance), g.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def history_sample(self):
        """The histories' values of the current state (pmx_history_sample)."""
        out = np.zeros(max(getattr(self, "_nhis", 0), 1))
        self._call("history_sample", self.h, out.ctypes.data_as(_dp))
        return out[:getattr(self, "_nhis", 0)]

    def upload(self, model):
        self._call("upload", self.h, C.byref(model.struct()))
        self._counts = (model.parnum, model.sprnum)

    def critical(self, which=CRIT_PARTICLE):
        """pmx_critical: per particle (CRIT_PARTICLE) or per spring
        (CRIT_SPRING) the critical step estimate, circular frequency and
        damping ratio, as three arrays."""
        n = self._counts[0] if which == CRIT_PARTICLE else self._counts[1]
        h, o, r = np.zeros(max(n, 1)), np.zeros(max(n, 1)), np.zeros(max(n, 1))
        self._call("critical", self.h, int(which), h.ctypes.data_as(_dp), o.ctypes.data_as(_dp),
                   r.ctypes.data_as(_dp))
        return h[:n], o[:n], r[:n]

    def dem_adaptive(self, duration, step, adaptive, forcing_=None):
        """pmx_dem_adaptive; forcing_ from forcing(on_step=...)."""
        el = C.c_double(0.0)
        try:
            self._call("dem_adaptive", self.h, duration, step, adaptive,
                       C.byref(forcing_) if forcing_ is not None else None, C.byref(el))
        except PmxError:
            errs = getattr(forcing_, "_step_errors", None)
            if errs:
                raise errs[0]
            raise
        return el.value

    def download(self, model):
        s = model.struct()
        self._call("download", self.h, C.byref(s))
        model.take_time(s)

    def dem(self, duration, step, forcing_=None):
        el = C.c_double(0.0)
        self._call("dem", self.h, duration, step, C.byref(forcing_) if forcing_ is not None else None, C.byref(el))
        return el.value

    def upload_batch(self, models):
        """pmx_upload_batch: the models are stepped together, each as if alone."""
        arr = (pmx_model * len(models))(*[m.struct() for m in models])
        self._call("upload_batch", self.h, arr, len(models))

    def download_batch(self, models):
        structs = [m.struct() for m in models]
        arr = (pmx_model * len(models))(*structs)
        self._call("download_batch", self.h, arr, len(models))
        for m, s in zip(models, arr):
            m.take_time(s)

    def dem_batch(self, duration, step, forcings=None):
        """pmx_dem_batch with one pmx_forcing per instance (or None)."""
        el = C.c_double(0.0)
        arr = None
        if forcings is not None:
            arr = (pmx_forcing * len(forcings))(*forcings)
            arr._keep = [getattr(f, "_keep", None) for f in forcings]
        self._call("dem_batch", self.h, duration, step, arr, C.byref(el))
        return el.value

    def contacts(self, instance=None):
        n = C.c_size_t(0)
        if instance is None:
            fn, pre = "contacts", ()
        else:
            fn, pre = "contacts_instance", (instance,)
        self._call(fn, self.h, *pre, None, 0, C.byref(n))
        arr = (pmx_pair * max(n.value, 1))()
        self._call(fn, self.h, *pre, arr, n.value, C.byref(n))
        out = np.zeros(n.value, dtype=[("master", np.int32), ("slave", np.int32), ("state", np.float64, 3)])
        if n.value:
            raw = np.ctypeslib.as_array(arr)[: n.value]
            out["master"] = raw["master"]
            out["slave"] = raw["slave"]
            out["state"] = raw["state"]
        return out

    def stats(self):
        s = pmx_stats()
        self._call("get_stats", self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in pmx_stats._fields_}

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.prefix + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The resident request loop (pmx_server_*): submit models with their
    duration, step and forcing; wait for each result."""

    def __init__(self, precision=PMX_FP64, device=0, row_capacity=0, max_batch=0, max_active=0, gather_s=0.0,
                 hbm_budget=0, repack_s=0.0, longest_first=False, lib=None):
        self.lib = lib or load()
        opt = pmx_server_options(precision, device, row_capacity, max_batch, max_active, gather_s, int(hbm_budget),
                                 float(repack_s), int(bool(longest_first)))
        h = C.c_void_p()
        self._check(self.lib.pmx_server_create(C.byref(opt), C.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != PMX_OK:
            raise PmxError(rc, self.lib.pmx_last_error().decode())

    def submit(self, model, duration, step, forcing_=None):
        i = C.c_longlong(0)
        self._check(self.lib.pmx_server_submit(self.h, C.byref(model.struct()), duration, step,
                                               C.byref(forcing_) if forcing_ is not None else None, C.byref(i)))
        return i.value

    def wait(self, rid, model=None, timeout=-1.0):
        """Blocks until request rid is done; fills model (a Model with the
        submitted counts) with its final state; returns the request info."""
        info = pmx_request_info()
        s = model.struct() if model is not None else None
        self._check(self.lib.pmx_server_wait(self.h, rid, timeout, C.byref(s) if s is not None else None,
                                             C.byref(info)))
        if model is not None:
            model.take_time(s)
        return {k: getattr(info, k) for k, _ in pmx_request_info._fields_}

    def stats(self):
        st = pmx_server_stats()
        self._check(self.lib.pmx_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in pmx_server_stats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.pmx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
