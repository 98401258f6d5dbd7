"""Parmec's Python input API (input.cpp:4093-4119) over the MI355X engine.

A deck written for Parmec runs unchanged against this module for the
supported subset:

    from parmec_amd.parmec import *
    mat = MATERIAL(2000.0, 1e6, 0.25)
    SPHERE((0.0, 0.0, 0.01), 0.005, mat, 1)
    OBSTACLE([(0,0,0, 1,0,0, 1,1,0)], 2)
    GRANULAR(0, 0, 1e6, 0.1, 0.3)
    GRAVITY(0.0, 0.0, -9.81)
    DEM(0.02, 1e-5)

Names, argument order, defaults and error types follow the reference:
MATERIAL (input.cpp:728-746), SPHERE (749-837), OBSTACLE (1367-1437),
GRANULAR (2384-2455), VELOCITY (2798-2837), GRAVITY (2840-2933),
DAMPING (2936-3013), TSERIES (626-725), DEM (3958-4091), RESET (610-623),
CRITICAL (3034-3154).  Gravity and damping callbacks and TSERIES are
evaluated on the host at each step's curtime (input.cpp:4424-4503) and
handed to the device as per-step tables, or step by step when DEM takes the
adaptive step.  The step loop itself is libpmx's pmx_dem / pmx_dem_adaptive
(C++ host + HIP kernels); CRITICAL is pmx_critical.

MESH (840-1156) builds a meshed particle from volume elements (its
surface triangles and mass characteristics as mesh.cpp:315-521 forms them;
particles of SPHERE and MESH share one numbering).  SPRING (1440-1788) is
supported with stroke offsets (load-curve TSERIES); TORSION_SPRING
(1791-2080), UNSPRING (2083-2295), BALL_JOINT (2298-2333) and RESTRAIN
(2458-2644) run on the device; OBSTACLE with point/linear/angular (moving
obstacles) and PRESCRIBE (2647-2795) tabulate their callables or TSERIES at
the step times.  Not on this path (raise NotImplementedError): ANALYTICAL,
XDMF and MED output (OUTPUT writes DUMP and VTK files), and HISTORY's
sphere/box sources and .h5 reading.
"""
import bisect
import math

import numpy as np

from . import pmx

__all__ = ["RESET", "TSERIES", "MATERIAL", "SPHERE", "OBSTACLE", "GRANULAR", "SPRING", "VELOCITY", "GRAVITY", "DAMPING",
           "DEM", "MESH", "ANALYTICAL", "PRESCRIBE", "HISTORY", "OUTPUT", "CRITICAL", "TORSION_SPRING", "UNSPRING",
           "RESTRAIN", "BALL_JOINT", "state", "contacts", "engine_options"]

_DENSITY, _YOUNG, _POISSON = 0, 1, 2

# output formats, modes and entities (constants.h:52-60)
OUT_FORMAT_DUMP, OUT_FORMAT_VTK, OUT_FORMAT_XDMF, OUT_FORMAT_MED = 1, 2, 4, 8
OUT_MODES = {"SPH": 1, "MESH": 2, "RB": 4, "CD": 8, "SL": 16, "ST": 32, "JT": 64}
OUT_ALL_MODES = 127
OUT_ENTITIES = {"NUMBER": 1, "COLOR": 2, "DISPL": 4, "ORIENT": 8, "ORIENT1": 16, "ORIENT2": 32, "ORIENT3": 64,
                "LINVEL": 128, "ANGVEL": 256, "FORCE": 512, "TORQUE": 1024, "F": 2048, "FN": 4096, "FT": 8192,
                "SF": 16384, "AREA": 32768, "PAIR": 65536, "LENGTH": 131072, "SS": 262144, "FF": 524288,
                "XDIR": 1048576, "YDIR": 2097152, "ZDIR": 4194304, "TRQROT": 8388608, "TRQTOT": 16777216,
                "TRQSPR": 33554432, "JREAC": 67108864}
OUT_DEFAULT_ENTITIES = sum(OUT_ENTITIES.values())   # every entity (parmec.cpp:1371-1373)
_OUTREST = 2   # particle flag: not listed by any OUTPUT subset (constants.h:39)


class _Series:
    """Piecewise-linear time series (timeseries.cpp:241-263)."""

    def __init__(self, times=None, values=None, constant=None):
        self.t = list(times or [])
        self.v = list(values or [])
        self.constant = constant

    def __call__(self, time):
        if self.constant is not None:
            return self.constant
        t, v = self.t, self.v
        if time < t[0]:
            return v[0]
        if time > t[-1]:
            return v[-1]
        i = min(max(bisect.bisect_right(t, time) - 1, 0), len(t) - 2)
        dt = time - t[i]
        return v[i] + (v[i + 1] - v[i]) * (dt / (t[i + 1] - t[i]))


class _State:
    def __init__(self):
        self.reset()

    def reset(self):
        self.materials = []          # (density, young, poisson)
        self.sph = []                # (cx, cy, cz, r, material, color, particle)
        self.parts = []              # particles in creation order: ("s", sphere) or ("m", mesh)
        self.meshes = []             # MESH particles (see _mesh_create)
        self.vel = {}                # particle -> (lin or None, ang or None)
        self.tris = []               # 9-tuples
        self.tricol = []
        self.triobs = []
        self.obstacles = []          # (first triangle, end, point, linear callable, angular callable)
        self.prescribed = []         # (particle, linear, angular, linear kind, angular kind)
        self.pairs = [(0, 0, [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.5])]   # pair_reset, parmec.cpp:375-388
        self.tms = []
        self.gravity = [0.0, 0.0, 0.0]
        self.gravfunc = [None, None, None]
        self.lindamp = None
        self.angdamp = None
        self.model = None            # SoA state once DEM has run
        self.engine = None
        self.dirty = True
        self.options = {"row_capacity": 0, "device": 0}
        self.springs = []            # SPRING records in creation order
        self.torsion = []            # TORSION_SPRING records in creation order
        self.unsprings = []          # UNSPRING records
        self.restraints = []         # RESTRAIN entries: (particle, linear [9], angular [9])
        self.joints = []             # BALL_JOINT entries: (part1, part2, point)
        self.lcurves = []            # load curves (x, y) in order of first use
        self.lcmap = {}              # TSERIES number -> load curve number (lcurve_from_time_series)
        self.sprmap = []             # spring number -> index of the sorted springs
        self.trqsprmap = []          # torsion spring number -> sorted index
        self.histories = []          # (entity, source list, point or None, the returned Python list)
        self.curtime_history = 0.0   # parmec.cpp:73, 2261
        self.outputs = []            # OUTPUT records: (particles, entities, modes)
        self.outrest = [OUT_DEFAULT_ENTITIES, OUT_ALL_MODES]   # parmec.cpp:1371-1375
        self.outformat = OUT_FORMAT_XDMF
        self.output_frame = 0
        self.curtime_output = 0.0    # parmec.cpp:72, 2260
        self.output_path = None


_S = _State()


def engine_options(**kw):
    """Device options of the engine (row_capacity, device)."""
    _S.options.update(kw)
    _S.engine = None


def _is_number(x):
    return isinstance(x, (int, float)) and not isinstance(x, bool)


def _tuple(x, name, n):
    if not isinstance(x, tuple) or len(x) != n:
        raise TypeError(f"'{name}' must be a tuple of {n} elements")
    return tuple(float(v) for v in x)


def RESET():
    """Reset the simulation (input.cpp:610-623)."""
    if _S.engine is not None:
        _S.engine.close()
    _S.reset()


def TSERIES(points):
    """Create a time series; returns its number (input.cpp:626-725)."""
    if _is_number(points):
        ts = _Series(constant=float(points))
    elif isinstance(points, str):
        data = np.loadtxt(points).reshape(-1, 2)
        ts = _Series(list(data[:, 0]), list(data[:, 1]))
    elif isinstance(points, list):
        if points and isinstance(points[0], (list, tuple)):
            ts = _Series([float(p[0]) for p in points], [float(p[1]) for p in points])
        else:
            if len(points) < 4:
                raise ValueError("Time series must have at least two points")
            ts = _Series([float(x) for x in points[0::2]], [float(x) for x in points[1::2]])
    else:
        raise TypeError("Invalid points format")
    _S.tms.append(ts)
    return len(_S.tms) - 1


def MATERIAL(density, young, poisson):
    """Create a bulk material; returns its number (input.cpp:728-746)."""
    for name, v in (("density", density), ("young", young), ("poisson", poisson)):
        if not v > 0:
            raise ValueError(f"'{name}' must be positive")
    _S.materials.append((float(density), float(young), float(poisson)))
    return len(_S.materials) - 1


def SPHERE(center, radius, material, color):
    """Create a spherical particle; returns the particle number (input.cpp:749-837)."""
    c = _tuple(center, "center", 3)
    if not radius > 0:
        raise ValueError("'radius' must be positive")
    if not (0 <= material < len(_S.materials)):
        raise ValueError("'material' must be >= 0 and < number of materials")
    if not color > 0:
        raise ValueError("'color' must be positive")
    _S.sph.append((c[0], c[1], c[2], float(radius), int(material), int(color), len(_S.parts)))
    _S.parts.append(("s", len(_S.sph) - 1))
    _S.dirty = True
    return len(_S.parts) - 1


def OBSTACLE(triangles, color, point=None, linear=None, angular=None):
    """Create obstacle triangles (input.cpp:1367-1437)."""
    if not isinstance(triangles, list) or any(not isinstance(t, tuple) or len(t) != 9 for t in triangles):
        raise TypeError("'triangles' must be a list of 9-element tuples")
    if isinstance(color, list):
        if len(color) != len(triangles):
            raise ValueError("'color' list must have one entry per triangle")
        cols = [int(c) for c in color]
    else:
        cols = [int(color)] * len(triangles)
    if any(c <= 0 for c in cols):
        raise ValueError("'color' must be positive")
    obs = -1
    if point is not None and (linear is not None or angular is not None):
        if not isinstance(point, tuple) or len(point) != 3:
            raise TypeError("'point' must be a 3-element tuple")
        for f, name in ((linear, "linear"), (angular, "angular")):
            if f is not None and not callable(f):
                raise TypeError(f"'{name}' must be callable")
        # the obstacle's own index k (triobs = -k-2, the value the contact law
        # decodes); the reference stores -(k+1)-2 here (input.cpp:1391, 1433),
        # which makes the law read the next obstacle's velocity
        obs = -len(_S.obstacles) - 2
        _S.obstacles.append((len(_S.tris), len(_S.tris) + len(triangles), tuple(float(x) for x in point), linear, angular))
    _S.tris.extend(tuple(float(x) for x in t) for t in triangles)
    _S.tricol.extend(cols)
    _S.triobs.extend([obs] * len(triangles))
    _S.dirty = True


# faces of the volume elements, in local node numbers 1..k (mesh.cpp:39-60);
# each face lists its nodes so that its normal points out of the element
_TET = ((1, 3, 2), (1, 2, 4), (2, 3, 4), (3, 1, 4))
_PYR = ((1, 4, 3, 2), (1, 2, 5), (2, 3, 5), (3, 4, 5), (4, 1, 5))
_WED = ((1, 3, 2), (4, 5, 6), (1, 2, 5, 4), (2, 3, 6, 5), (3, 1, 4, 6))
_HEX = ((1, 4, 3, 2), (1, 2, 6, 5), (2, 3, 7, 6), (3, 4, 8, 7), (1, 5, 8, 4), (5, 6, 7, 8))
_ELEMENT_FACES = {4: _TET, 5: _PYR, 6: _WED, 8: _HEX}


def _simplex(J, a, b, c, d):
    """Integrals over a tetrahedron of 1, x, y, z and the products of the
    coordinates, scaled by J (the density times the Jacobian): simplex.h:33-120,
    (J/6, J/24 sum x, J/120 (sum x_i y_i + sum_i sum_j x_i y_j))."""
    p = (a, b, c, d)

    def second(u, v):
        t = 0.0
        for i in range(4):
            for j in range(4):
                t = t + (2.0 * p[i][u] * p[j][v] if i == j else p[i][u] * p[j][v])
        return (J / 120.0) * t
    first = [(J / 24.0) * (a[k] + b[k] + c[k] + d[k]) for k in range(3)]
    return J / 6.0, first, (second(0, 0), second(0, 1), second(1, 1), second(0, 2), second(1, 2), second(2, 2))


def _jacobian(a, b, c, d):
    q = [b[0] - a[0], c[0] - a[0], d[0] - a[0], b[1] - a[1], c[1] - a[1], d[1] - a[1], b[2] - a[2], c[2] - a[2],
         d[2] - a[2]]
    return (q[0] * q[4] * q[8] + q[3] * q[7] * q[2] + q[6] * q[1] * q[5] - q[6] * q[4] * q[2] - q[0] * q[7] * q[5] -
            q[3] * q[1] * q[8])


def _mesh_create(xyz, elements, colors):
    """Surface triangles and mass characteristics of a volume mesh
    (MESH_Create and MESH_Char, mesh.cpp:315-521).

    elements: [(node list, material)], colors: (global colour, {sorted face
    nodes: colour}).  A face shared by two elements is inner; the rest are
    the surface.  The surface faces are listed in the order the reference
    leaves them: the surface elements in input order, each element's faces
    in its face-table order, the whole list reversed; a quadrilateral gives
    the triangles (0, 1, 2) and (0, 2, 3).  Node numbers become local to the
    range of nodes the elements use.  Mass, centre and inertia integrate
    each element as tetrahedra fanned from the range's first node over
    fan-triangulated faces (bulk elements first, then surface elements, each
    in input order): sums of density-weighted monomials, shifted to the
    centre, inertia = trace(E) I - E of the Euler tensor E."""
    gcolor, fcolor = colors
    lo = min(min(nodes) for nodes, _ in elements)
    hi = max(max(nodes) for nodes, _ in elements)
    # faces: the first element to list a face owns it; a later one makes it inner
    owner = {}
    shared = [0] * len(elements)
    order = []
    for e, (nodes, _) in enumerate(elements):
        for n, f in enumerate(_ELEMENT_FACES[len(nodes)]):
            key = tuple(sorted(nodes[t - 1] for t in f))
            if key in owner:
                shared[e] += 1
                o = owner[key]
                if o is not None:
                    shared[o[0]] += 1
                owner[key] = None
            else:
                owner[key] = (e, n)
                order.append(key)
    surface = [e for e in range(len(elements)) if shared[e] < len(_ELEMENT_FACES[len(elements[e][0])])]
    bulk = [e for e in range(len(elements)) if e not in set(surface)]
    faces_of = {}
    for key in order:
        o = owner[key]
        if o is not None:
            faces_of.setdefault(o[0], []).append((o[1], key))
    faces = []
    for e in surface:
        nodes = elements[e][0]
        for n, key in sorted(faces_of.get(e, [])):
            loc = [nodes[t - 1] - lo for t in _ELEMENT_FACES[len(nodes)][n]]
            faces.append((loc, fcolor.get(key, gcolor)))
    faces.reverse()
    tris = []
    for loc, col in faces:
        tris.append(((loc[0], loc[1], loc[2]), col))
        if len(loc) == 4:
            tris.append(((loc[0], loc[2], loc[3]), col))
    local = xyz[lo:hi + 1]
    zero = tuple(local[0])
    me = 0.0
    sxyz = [0.0, 0.0, 0.0]
    E = [0.0] * 9
    for e in bulk + surface:
        nodes, rho = elements[e]
        for f in _ELEMENT_FACES[len(nodes)]:
            a = tuple(local[nodes[f[0] - 1] - lo])
            for j in range(1, len(f) - 1):
                b = tuple(local[nodes[f[j] - 1] - lo])
                c = tuple(local[nodes[f[j + 1] - 1] - lo])
                J = rho * _jacobian(zero, a, b, c)
                v, first, sec = _simplex(J, zero, a, b, c)
                me += v
                sxyz = [sxyz[k] + first[k] for k in range(3)]
                E[0] += sec[0]
                E[3] += sec[1]
                E[4] += sec[2]
                E[6] += sec[3]
                E[7] += sec[4]
                E[8] += sec[5]
    sx, sy, sz = sxyz
    cen = (sx / me, sy / me, sz / me)
    E[0] -= (2 * sx - cen[0] * me) * cen[0]
    E[4] -= (2 * sy - cen[1] * me) * cen[1]
    E[8] -= (2 * sz - cen[2] * me) * cen[2]
    E[3] -= cen[0] * sy + cen[1] * sx - cen[0] * cen[1] * me
    E[6] -= cen[0] * sz + cen[2] * sx - cen[0] * cen[2] * me
    E[7] -= cen[1] * sz + cen[2] * sy - cen[1] * cen[2] * me
    E[1], E[2], E[5] = E[3], E[6], E[7]
    tr = E[0] + E[4] + E[8]
    inertia = [(tr if k in (0, 4, 8) else 0.0) - E[k] for k in range(9)]
    return {"nodes": [tuple(x) for x in local], "triangles": tris, "mass": me, "center": cen, "inertia": inertia,
            "elements": [([n - lo for n in nodes], vol) for nodes, vol in elements]}


def MESH(nodes, elements, material, colors):
    """Create a meshed particle; returns the particle number (input.cpp:840-1156).

    nodes: [x0, y0, z0, x1, ...]; elements: [k, n_1 .. n_k, volid, ...] with
    k = 4, 5, 6 or 8 (tetrahedron, pyramid, wedge, hexahedron) and volid the
    material whose density the element has; colors: the surface colour, or
    [colour, k, n_1 .. n_k, colour, ...] with k = 3 or 4 face colours over
    the global one.  The surface faces become triangles owned by the
    particle (triobs = its number)."""
    if not isinstance(nodes, list) or not isinstance(elements, list):
        raise TypeError("'nodes' and 'elements' must be lists")
    if not (isinstance(colors, list) or _is_number(colors)):
        raise TypeError("'colors' must be a list or a number")
    if not (0 <= material < len(_S.materials)):
        raise ValueError("'material' must be >= 0 and < number of materials")
    els, i, l = [], 0, len(elements)
    while i < l:
        k = int(elements[i])
        if k not in (4, 5, 6, 8):
            raise ValueError("An element must have 4, 5, 6, or 8 nodes")
        if i + k + 1 >= l:
            raise ValueError("The last element definition is incomplete")
        els.append(([int(x) for x in elements[i + 1:i + 1 + k]], int(elements[i + 1 + k])))
        i += k + 2
    if len(nodes) % 3:
        raise ValueError("Nodes list length must be a multiple of 3")
    nn = len(nodes) // 3
    for m_, (ns, vol) in enumerate(els):
        for j, n in enumerate(ns):
            if n < 0 or n >= nn:
                raise ValueError(f"Node {j} in element {m_} is outside of range [0, {nn - 1}]")
        if len(set(ns)) != len(ns):
            raise ValueError(f"Element {m_} repeats a node")
        if not (0 <= vol < len(_S.materials)):
            raise ValueError(f"Element {m_} material (volid) {vol} is not a material number")
    if not els:
        raise ValueError("A mesh needs at least one element")
    fcolor = {}
    if isinstance(colors, list):
        gcolor = int(colors[0])
        i, l = 1, len(colors)
        while i < l:
            k = int(colors[i])
            if k not in (3, 4):
                raise ValueError("A face must have 3 or 4 nodes")
            if i + k + 1 >= l:
                raise ValueError("The last face definition is incomplete")
            fn = [int(x) for x in colors[i + 1:i + 1 + k]]
            for j, n in enumerate(fn):
                if n < 0 or n >= nn:
                    raise ValueError(f"Node {j} in face {len(fcolor)} is outside of range [0, {nn - 1}]")
            if len(set(fn)) != len(fn):
                raise ValueError(f"Face {len(fcolor)} repeats a node")
            fcolor[tuple(sorted(fn))] = int(colors[i + 1 + k])
            i += k + 2
    else:
        gcolor = int(colors)
    xyz = [tuple(float(nodes[3 * q + c]) for c in range(3)) for q in range(nn)]
    dens = [_S.materials[vol][_DENSITY] for _, vol in els]
    msh = _mesh_create(xyz, [(ns, dens[e]) for e, (ns, _) in enumerate(els)], (gcolor, fcolor))
    part = len(_S.parts)
    msh["part"] = part
    msh["material"] = int(material)
    msh["first_triangle"] = len(_S.tris)
    for (a, b, c), col in msh["triangles"]:
        x = msh["nodes"]
        _S.tris.append(tuple(x[a]) + tuple(x[b]) + tuple(x[c]))
        _S.tricol.append(col)
        _S.triobs.append(part)
    _S.meshes.append(msh)
    _S.parts.append(("m", len(_S.meshes) - 1))
    _S.dirty = True
    return part


def GRANULAR(color1, color2, spring, damper=1.0, friction=0.0, rolling=0.0, drilling=0.0, kskn=0.5):
    """Granular interaction parameters of a colour pair (input.cpp:2384-2455)."""
    if not spring > 0:
        raise ValueError("'spring' must be positive")
    if damper < 0 or rolling < 0 or drilling < 0:
        raise ValueError("'damper', 'rolling' and 'drilling' must be non-negative")
    if not kskn > 0:
        raise ValueError("'kskn' must be positive")
    if isinstance(friction, tuple):
        if len(friction) != 2:
            raise TypeError("'friction' must be a number or a (static, dynamic) tuple")
        fri = (float(friction[0]), float(friction[1]))
    else:
        fri = (float(friction), float(friction))
    if fri[0] < 0 or fri[1] < 0:
        raise ValueError("'friction' must be non-negative")
    par = [float(spring), float(damper), fri[0], fri[1], float(rolling), float(drilling), float(kskn)]
    if color1 == 0 and color2 == 0:
        _S.pairs[0] = (0, 0, par)
    elif color1 > 0 and color2 > 0:
        _S.pairs.append((min(color1, color2), max(color1, color2), par))
    else:
        raise ValueError("Invalid color values")
    _S.dirty = True


def VELOCITY(parnum, linear=None, angular=None):
    """Set initial velocities of a particle (input.cpp:2798-2837)."""
    if not (0 <= parnum < len(_S.parts)):
        raise ValueError("'parnum' must be >= 0 and < number of particles")
    lin = _tuple(linear, "linear", 3) if linear is not None else None
    ang = _tuple(angular, "angular", 3) if angular is not None else None
    if _S.model is not None:
        m = _S.model
        if lin:
            m.linear[:, parnum] = lin
        if ang:
            m.angular[3:, parnum] = ang
            L = m.rotation[:, parnum]
            o = ang
            m.angular[0, parnum] = L[0] * o[0] + L[1] * o[1] + L[2] * o[2]   # TVMUL, macros.h:966-971
            m.angular[1, parnum] = L[3] * o[0] + L[4] * o[1] + L[5] * o[2]
            m.angular[2, parnum] = L[6] * o[0] + L[7] * o[1] + L[8] * o[2]
        _S.dirty = True
    else:
        _S.vel[parnum] = (lin, ang)


def _component(x, name):
    if callable(x):
        return ("func", x)
    if isinstance(x, float):
        return ("const", x)
    if isinstance(x, int) and not isinstance(x, bool):
        if not (0 <= x < len(_S.tms)):
            raise ValueError(f"Invalid TSERIES number for {name}")
        return ("tms", x)
    raise ValueError(f"Invalid {name} component")


def GRAVITY(gx, gy, gz):
    """Set gravity; a float is constant, a callable is g(t), an int is a TSERIES
    number (input.cpp:2840-2933)."""
    comps = [_component(g, n) for g, n in ((gx, "gx"), (gy, "gy"), (gz, "gz"))]
    for k, (kind, v) in enumerate(comps):
        if kind == "const":
            _S.gravity[k] = v
            _S.gravfunc[k] = None
        elif kind == "func":
            _S.gravfunc[k] = v
        else:
            _S.gravfunc[k] = _S.tms[v]


def DAMPING(linear, angular):
    """Global damping: callables returning 3-tuples, or 3-tuples of TSERIES
    numbers (input.cpp:2936-3013)."""
    def one(x, name):
        if callable(x):
            return x
        if isinstance(x, tuple):
            if len(x) != 3:
                raise ValueError(f"Invalid {name} damping tuple size")
            ids = [int(v) for v in x]
            if any(not (0 <= i < len(_S.tms)) for i in ids):
                raise ValueError(f"Invalid TSERIES number in {name} damping tuple")
            return lambda t, ids=ids: tuple(_S.tms[i](t) for i in ids)
        raise ValueError(f"Invalid {name} damping argument type")
    _S.lindamp = one(linear, "linear")
    _S.angdamp = one(angular, "angular")


_REAL_MAX = 1.7976931348623157e308


def _norm3(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return n


def SPRING(part1, point1, part2, geom2, spring, dashpot=None, direction=None, planar=None, unload=None, ylim=None,
           inactive=False, offset=-1, friction=0.0, kskn=0.0):
    """Translational spring between particles (input.cpp:1440-1788); returns
    the spring number.  geom2 is a point (part2 >= 0, or -1 for a fixed
    point) or [point, normal] for a point-to-plane spring."""
    if not (isinstance(part1, int) and part1 >= 0):
        raise ValueError("'part1' must be non-negative")
    p1 = _tuple(point1, "point1", 3)
    if offset < -1 or offset >= len(_S.tms):
        raise ValueError("offset TSERIES index is out of range")
    if part2 < -1:
        raise ValueError("Particle two index is out of range [-1, 0, ...]")
    if not isinstance(spring, list) or len(spring) < 4 or len(spring) % 2:
        raise ValueError("Invalid spring lookup table list length")
    if isinstance(dashpot, list) and (len(dashpot) < 4 or len(dashpot) % 2):
        raise ValueError("Invalid dashpot lookup table list length")
    if unload is not None and (len(unload) < 4 or len(unload) % 2):
        raise ValueError("Invalid unload lookup table list length")
    if friction < 0 or kskn < 0:
        raise ValueError("'friction' and 'kskn' must be non-negative")
    r = {"part": (part1, part2), "pnt": [list(p1), None], "dir": [0.0] * 6, "flg": 0, "fric": float(friction),
         "kskn": float(kskn), "unspring": -1 if inactive else -3, "yield": (0.0, 0.0), "unload": None,
         "offset": -1 if offset < 0 else _lcurve(offset)}
    if isinstance(geom2, tuple):
        r["pnt"][1] = list(_tuple(geom2, "geom2", 3))
        project = False
    elif isinstance(geom2, list) and len(geom2) == 2:
        q = list(_tuple(geom2[0], "geom2", 3))
        d = list(_tuple(geom2[1], "geom2", 3))
        ln = _norm3(d)
        d = [d[0] / ln, d[1] / ln, d[2] / ln]            # NORMALIZE, macros.h:427-433
        r["dir"] = d + d
        w = [p1[0] - q[0], p1[1] - q[1], p1[2] - q[2]]
        ln = d[0] * w[0] + d[1] * w[1] + d[2] * w[2]
        r["pnt"][1] = [p1[0] - ln * d[0], p1[1] - ln * d[1], p1[2] - ln * d[2]]   # projection onto the plane
        project = True
    else:
        raise TypeError("'geom2' must be a tuple or a [point, normal] list")
    xs = [float(v) for v in spring[0::2]]
    if any(xs[k] <= xs[k - 1] for k in range(1, len(xs))):
        raise ValueError("Spring stroke values must increase")
    r["spring"] = (xs, [float(v) for v in spring[1::2]])
    if isinstance(dashpot, list):
        vs = [float(v) for v in dashpot[0::2]]
        if any(vs[k] <= vs[k - 1] for k in range(1, len(vs))):
            raise ValueError("Dashpot velocity values must increase")
        r["dashpot"] = (vs, [float(v) for v in dashpot[1::2]])
    elif dashpot is not None:
        ratio = float(dashpot)
        if ratio < 0.0:
            raise ValueError("Critical damping ratio not in [0.0, +Inf) interval")
        r["dashpot"] = ([ratio], [ratio])
    else:
        r["dashpot"] = ([-_REAL_MAX, _REAL_MAX], [0.0, 0.0])
    if unload is not None:
        us, uf = [float(v) for v in unload[0::2]], [float(v) for v in unload[1::2]]
        for k in range(1, len(us)):
            if (uf[k] - uf[k - 1]) / (us[k] - us[k - 1]) <= 0.0:
                raise ValueError("Unloading curve must be monotonically increasing")
        r["unload"] = (us, uf)
    p2 = r["pnt"][1]
    if project:
        r["flg"] |= pmx.SPRDIR_PROJECT
        r["stroke0"] = 0.0
    elif direction is not None:
        d = list(_tuple(direction, "direction", 3))
        ln = _norm3(d)
        if ln == 0.0:
            raise ValueError("Invalid zero direction")
        inv = 1.0 / ln
        d = [d[0] * inv, d[1] * inv, d[2] * inv]
        r["dir"] = d + d
        if planar is None or planar == "OFF":
            r["flg"] |= pmx.SPRDIR_CONSTANT
        elif planar == "ON":
            r["flg"] |= pmx.SPRDIR_PLANAR
        else:
            raise ValueError("Invalid planar switch")
        dif = [p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]]
        dot = dif[0] * d[0] + dif[1] * d[1] + dif[2] * d[2]
        if r["flg"] & pmx.SPRDIR_CONSTANT:
            r["stroke0"] = dot
        else:
            dif = [dif[0] - dot * d[0], dif[1] - dot * d[1], dif[2] - dot * d[2]]
            r["stroke0"] = _norm3(dif)
    else:
        r["flg"] |= pmx.SPRDIR_FOLLOWER
        dif = [p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]]
        r["stroke0"] = _norm3(dif)
    if ylim is not None:
        y = _tuple(ylim, "ylim", 2)
        if y[0] > 0.0:
            raise ValueError("Compressive yield limit must be <= 0.0")
        if y[1] < 0.0:
            raise ValueError("Tensile yield limit must be >= 0.0")
        r["yield"] = y
    _S.springs.append(r)
    _S.dirty = True
    return len(_S.springs) - 1


def _lcurve(ts):
    """The load curve of a TSERIES (lcurve_from_time_series,
    parmec.cpp:1239-1269): its points as a look-up table, made once per
    series."""
    if ts not in _S.lcmap:
        series = _S.tms[ts]
        if series.constant is not None or len(series.t) < 2:
            raise ValueError("a load curve TSERIES needs at least two points")
        _S.lcmap[ts] = len(_S.lcurves)
        _S.lcurves.append((list(series.t), list(series.v)))
    return _S.lcmap[ts]


def _curve_table(x, name, what):
    """A look-up table list [x0, y0, x1, y1, ...] with increasing x."""
    if len(x) < 4 or len(x) % 2:
        raise ValueError(f"Invalid {name} lookup table list length")
    xs, ys = [float(v) for v in x[0::2]], [float(v) for v in x[1::2]]
    if any(xs[k] <= xs[k - 1] for k in range(1, len(xs))):
        raise ValueError(what)
    return xs, ys


def TORSION_SPRING(part1, part2, zdir, xdir, kroll=None, kpitch=None, kyaw=None, droll=None, dpitch=None, dyaw=None,
                   cone=None, refpnt=None):
    """Torsion spring between two particles, or a particle and the ground
    (part2 = -1), about the roll (zdir), pitch and yaw axes of part1's frame
    (input.cpp:1791-2080); returns the torsion spring number.  kroll, kpitch,
    kyaw: [angle, torque, ...] curves (default zero torque); droll, dpitch,
    dyaw: [angular velocity, torque, ...] curves or a critical damping ratio;
    cone: two or three of 'roll', 'pitch', 'yaw' whose angles act as one.
    As the reference does, zdir is scaled by 1/|zdir|^2 and xdir is kept as
    given (input.cpp:1882-1888)."""
    n = len(_S.parts)
    if not (isinstance(part1, int) and 0 <= part1 < n):
        raise ValueError("'part1' must be >= 0 and < the number of particles")
    if not isinstance(part2, int) or part2 < -1:
        raise ValueError("Particle two index is out of range [-1, 0, ...]")
    if part2 >= n:
        raise ValueError("Particle two index is out of range")
    z = list(_tuple(zdir, "zdir", 3))
    x = list(_tuple(xdir, "xdir", 3))
    curves = []
    for name, c in (("kroll", kroll), ("kpitch", kpitch), ("kyaw", kyaw)):
        if c is not None and not isinstance(c, list):
            raise TypeError(f"'{name}' must be a list")
        curves.append(_curve_table(c, name, "Torsion spring angle values must increase") if c is not None
                      else ([-_REAL_MAX, _REAL_MAX], [0.0, 0.0]))
    dampers = []
    for name, d in (("droll", droll), ("dpitch", dpitch), ("dyaw", dyaw)):
        if isinstance(d, list):
            dampers.append(_curve_table(d, name, "Torsion damper angular velocity values must increase"))
        elif d is not None:
            if not _is_number(d):
                raise TypeError(f"'{name}' must be a list or a number")
            if d < 0.0:
                raise ValueError("Critical damping ratio not in [0.0, +Inf) interval")
            dampers.append(([float(d)], [float(d)]))
        else:
            dampers.append(([-_REAL_MAX, _REAL_MAX], [0.0, 0.0]))
    zlen0 = _norm3(z)
    xlen0 = _norm3(z)   # the reference measures zdir twice (input.cpp:1878)
    if zlen0 == 0.0:
        raise ValueError("Zero zdir length")
    if xlen0 == 0.0:
        raise ValueError("Zero xdir length")
    izl0, ixl0 = 1.0 / zlen0, 1.0 / xlen0
    z = [v * izl0 for v in z]
    z = [v * ixl0 for v in z]
    if z[0] * x[0] + z[1] * x[1] + z[2] * x[2] > 1e-8 * 1e-8:
        raise ValueError("Directions zdir and xdir do not seem orthogonal")
    flags = 0
    if cone is not None:
        if not isinstance(cone, tuple):
            raise TypeError("'cone' must be a tuple")
        if not 2 <= len(cone) <= 3:
            raise ValueError("Invalid cone tuple size: neither 2 nor 3")
        for a in cone:
            if not isinstance(a, str):
                raise ValueError("Invalid cone tuple component: not a string")
            if a not in ("roll", "pitch", "yaw"):
                raise ValueError("Invalid cone tuple component name")
            flags |= {"roll": pmx.TRQCONE_ROLL, "pitch": pmx.TRQCONE_PITCH, "yaw": pmx.TRQCONE_YAW}[a]
        if flags in (3, 5, 7) and kroll is None and droll is None:
            raise ValueError("A roll curve not defined for cone = " + str(cone))
        if flags == 6 and kpitch is None and dpitch is None:
            raise ValueError("A pitch curve not defined for cone = ('pitch', 'yaw')")
        if flags not in (3, 5, 6, 7):
            raise ValueError("Invalid cone tuple: components repeat")
    ref = _tuple(refpnt, "refpnt", 3) if refpnt is not None else (_REAL_MAX,) * 3
    _S.torsion.append({"part": (part1, part2), "z": z, "x": x, "k": curves, "d": dampers, "cone": flags,
                       "refpnt": ref})
    _S.dirty = True
    return len(_S.torsion) - 1


def UNSPRING(tsprings, msprings, limits, entity=None, operator=None, abs=False, nsteps=1, nfreq=1, unload=-1,
             activate=None):
    """Switch springs off when an entity of other springs leaves its limits
    (input.cpp:2083-2295): on every nsteps-th DEM call, the SUM (default),
    MIN or MAX of STROKE, F, SF (default) or FF over tsprings (|values| with
    abs) is tested against limits = (lower or None, upper or None); after
    nfreq failures the msprings drop to zero force, or unload along TSERIES
    `unload`, and the inactive springs in `activate` switch on.  Returns the
    condition's number."""
    if not isinstance(tsprings, list) or not isinstance(msprings, list):
        raise TypeError("'tsprings' and 'msprings' must be lists")
    if not isinstance(limits, tuple) or len(limits) != 2:
        raise TypeError("'limits' must be a tuple of 2 elements")
    if activate is not None and not isinstance(activate, list):
        raise TypeError("'activate' must be a list")
    if not isinstance(abs, bool):
        raise TypeError("'abs' must be a boolean")
    if not (isinstance(nsteps, int) and nsteps > 0) or not (isinstance(nfreq, int) and nfreq > 0):
        raise ValueError("'nsteps' and 'nfreq' must be positive")
    if unload < -1 or unload >= len(_S.tms):
        raise ValueError("unload TSERIES index is out of range")
    if unload >= 0:
        ts = _S.tms[unload]
        for j in range(1, len(ts.t)):
            if (ts.v[j] - ts.v[j - 1]) / (ts.t[j] - ts.t[j - 1]) <= 0.0:
                raise ValueError("Unloading curve must be monotonically increasing")
    S = len(_S.springs)
    for j in tsprings:
        if j < 0 or j >= S:
            raise ValueError("tsprings SPRING index is out of range")
    for j in msprings:
        if j < 0 or j >= S:
            raise ValueError("msprings SPRING index is out of range")
        if _S.springs[j]["unspring"] == -2:
            raise ValueError(f"in msprings, spring index {j} already in use by an earlier UNSPRING call")
        if _S.springs[j]["unspring"] == -1:
            raise ValueError(f"in msprings, trying to use an inactive SPRING with index {j}")
    lo = -_REAL_MAX if limits[0] is None else limits[0]
    hi = _REAL_MAX if limits[1] is None else limits[1]
    if not _is_number(lo):
        raise ValueError("Invalid lower limit")
    if not _is_number(hi):
        raise ValueError("Invalid upper limit")
    if lo >= hi:
        raise ValueError("Lower limit >= upper limit")
    ents = {"STROKE": pmx.HIS_ENTITIES.index("STROKE"), "F": pmx.HIS_ENTITIES.index("F"),
            "SF": pmx.HIS_ENTITIES.index("SF"), "FF": pmx.HIS_ENTITIES.index("FF")}
    if entity is not None and entity not in ents:
        raise ValueError("Invalid entity")
    ops = {"SUM": pmx.OP_SUM, "MIN": pmx.OP_MIN, "MAX": pmx.OP_MAX}
    if operator is not None and operator not in ops:
        raise ValueError("Invalid operator")
    for j in activate or []:
        if j < 0 or j >= S or _S.springs[j]["unspring"] != -1:
            raise ValueError(f"in activate list: SPRING {j} is already active")
    for j in msprings:
        _S.springs[j]["unspring"] = -2   # used by UNSPRING (input.cpp:2184)
    _S.unsprings.append({"t": [int(j) for j in tsprings], "m": [int(j) for j in msprings],
                         "a": [int(j) for j in activate or []], "lim": (float(lo), float(hi)),
                         "ent": ents[entity or "SF"], "op": ops[operator or "SUM"], "abs": 1 if abs else 0,
                         "nsteps": nsteps, "nfreq": nfreq, "action": -1 if unload < 0 else _lcurve(unload)})
    _S.dirty = True
    return len(_S.unsprings) - 1


def _directions(d, name):
    """Up to three orthonormal directions from 3, 6 or 9 numbers
    (input.cpp:2486-2559: normalised, then Gram-Schmidt)."""
    def normalize(v):
        ln = _norm3(v)
        return [v[0] / ln, v[1] / ln, v[2] / ln]

    def submul(a, eps, b):
        return [a[0] - eps * b[0], a[1] - eps * b[1], a[2] - eps * b[2]]

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    v = [float(x) for x in d]
    z = [0.0, 0.0, 0.0]
    if len(v) == 3:
        return normalize(v[0:3]) + z + z
    if len(v) == 6:
        d0 = normalize(v[0:3])
        d1 = normalize(submul(v[3:6], dot(d0, v[3:6]), d0))
        return d0 + d1 + z
    if len(v) == 9:
        d0 = normalize(v[0:3])
        d1 = normalize(submul(v[3:6], dot(d0, v[3:6]), d0))
        d2 = submul(v[6:9], dot(d0, v[6:9]), d0)
        d2 = normalize(submul(d2, dot(d1, d2), d1))
        return d0 + d1 + d2
    raise ValueError(f"Invalid {name} constrained directions list")


def RESTRAIN(parnum, linear=None, angular=None):
    """Restrain a particle's motion along up to three linear and three
    angular directions (input.cpp:2458-2644): [x, y, z], [x0, y0, z0, x1,
    y1, z1] or nine numbers, orthonormalised."""
    if not (isinstance(parnum, int) and 0 <= parnum < len(_S.parts)):
        raise ValueError("'parnum' must be >= 0 and < the number of particles")
    for name, x in (("linear", linear), ("angular", angular)):
        if x is not None and not isinstance(x, list):
            raise TypeError(f"'{name}' must be a list")
    if linear is None and angular is None:
        return None
    lin = _directions(linear, "linear") if linear is not None else [0.0] * 9
    ang = _directions(angular, "angular") if angular is not None else [0.0] * 9
    _S.restraints.append((parnum, lin, ang))
    _S.dirty = True
    return None


def BALL_JOINT(part1, point, part2=-1):
    """Algebraic ball joint of part1 at `point` to part2 or to the ground
    (part2 = -1) (input.cpp:2298-2333); returns the joint number."""
    n = len(_S.parts)
    if not (isinstance(part1, int) and 0 <= part1 < n):
        raise ValueError("'part1' must be >= 0 and < the number of particles")
    pt = _tuple(point, "point", 3)
    if part2 >= 0 and part2 >= n:
        raise ValueError("Particle number 'part2' is out of range")
    _S.joints.append((part1, int(part2) if part2 >= 0 else -1, pt))
    _S.dirty = True
    return len(_S.joints) - 1


def _f3_arrays(m):
    """Load curves, torsion springs (sorted by part1, part2; ties keep
    creation order), UNSPRING conditions (spring numbers mapped to the
    sorted springs), restraints and joints into the model."""
    m.set_lcurves(_S.lcurves)
    recs = _S.torsion
    order = sorted(range(len(recs)), key=lambda i: (recs[i]["part"][0], recs[i]["part"][1], i))
    T = len(recs)
    _S.trqsprmap = [0] * T
    nk = sum(len(recs[i]["k"][o][0]) for i in order for o in range(3))
    nd = sum(len(recs[i]["d"][o][0]) for i in order for o in range(3))
    old = _S.model
    m.set_torsion(T, nk, nd)
    ck = cd = 0
    for o in range(3):
        for k, i in enumerate(order):
            r = recs[i]
            _S.trqsprmap[i] = k
            x, y = r["k"][o]
            m.krpyidx[o, k] = ck
            m.krpy[0, ck:ck + len(x)], m.krpy[1, ck:ck + len(x)] = x, y
            ck += len(x)
            x, y = r["d"][o]
            m.drpyidx[o, k] = cd
            m.drpy[0, cd:cd + len(x)], m.drpy[1, cd:cd + len(x)] = x, y
            cd += len(x)
        m.krpyidx[o, T], m.drpyidx[o, T] = ck, cd
    for k, i in enumerate(order):
        r = recs[i]
        m.trqsprpart[:, k] = r["part"]
        m.trqzdir0[:, k], m.trqxdir0[:, k] = r["z"], r["x"]
        m.trqcone[k] = r["cone"]
    if old is not None and old.trqsprnum == T and T:
        for name in ("trqzdir1", "trqxdir1", "trqrpy", "trqrpytot", "trqrpyspr"):
            getattr(m, name)[:] = getattr(old, name)
    U = _S.unsprings
    m.set_unsprings(len(U), sum(len(u["t"]) for u in U), sum(len(u["m"]) for u in U), sum(len(u["a"]) for u in U))
    at = am = aa = 0
    for q, u in enumerate(U):
        m.tspridx[q], m.mspridx[q], m.actidx[q] = at, am, aa
        for j in u["t"]:
            m.tsprings[at] = _S.sprmap[j]
            at += 1
        for j in u["m"]:
            m.msprings[am] = _S.sprmap[j]
            am += 1
        for j in u["a"]:
            m.activate[aa] = _S.sprmap[j]
            aa += 1
        m.unlim[:, q] = u["lim"]
        m.unent[q], m.unop[q], m.unabs[q] = u["ent"], u["op"], u["abs"]
        m.nsteps[q], m.nfreq[q], m.unaction[q] = u["nsteps"], u["nfreq"], u["action"]
    n = len(U)
    m.tspridx[n], m.mspridx[n], m.actidx[n] = at, am, aa
    if old is not None and old.unsprnum == n and n:
        m.nfreq[:] = old.nfreq
    m.set_restraints(len(_S.restraints))
    for q, (p, lin, ang) in enumerate(_S.restraints):
        m.rstpart[q] = p
        m.rstlin[:, q], m.rstang[:, q] = lin, ang
    m.set_joints(len(_S.joints))
    for q, (p1, p2, pt) in enumerate(_S.joints):
        m.jpart[:, q] = (p1, p2)
        m.jpoint[:, q] = pt


def _spring_arrays(m):
    """Spring records -> SoA arrays, sorted by (part1, part2) as sort_springs
    leaves them (parmec.cpp:1683-1966; ties keep creation order)."""
    recs = _S.springs
    order = sorted(range(len(recs)), key=lambda i: (recs[i]["part"][0], recs[i]["part"][1], i))
    S = len(recs)
    _S.sprmap = [0] * S
    for k, i in enumerate(order):
        _S.sprmap[i] = k
    ncrv = sum(len(recs[i]["spring"][0]) for i in order)
    ndash = sum(len(recs[i]["dashpot"][0]) for i in order)
    nun = sum(len(recs[i]["unload"][0]) for i in order if recs[i]["unload"])
    m.set_springs(S, ncrv, ndash, nun)
    cs = ds = us = 0
    for k, i in enumerate(order):
        r = recs[i]
        m.sprpart[:, k] = r["part"]
        for p in range(2):
            m.sprpnt[p, :3, k] = r["pnt"][p]
            m.sprpnt[p, 3:, k] = r["pnt"][p]
        m.sprdir[:, k] = r["dir"]
        m.sprflg[k] = r["flg"]
        m.sprfric[k], m.sprkskn[k] = r["fric"], r["kskn"]
        m.unspring[k] = r["unspring"]
        m.sproffset[k] = r["offset"]
        m.yield_[:, k] = r["yield"]
        m.stroke0[k] = r["stroke0"]
        x, y = r["spring"]
        m.spridx[k] = cs
        m.spring[0, cs:cs + len(x)], m.spring[1, cs:cs + len(x)] = x, y
        cs += len(x)
        x, y = r["dashpot"]
        m.dashidx[k] = ds
        m.dashpot[0, ds:ds + len(x)], m.dashpot[1, ds:ds + len(x)] = x, y
        ds += len(x)
        m.unidx[k] = us
        if r["unload"]:
            x, y = r["unload"]
            m.unload[0, us:us + len(x)], m.unload[1, us:us + len(x)] = x, y
            us += len(x)
            m.sprtype[k] = pmx.SPRING_GENERAL_NONLINEAR
        else:
            m.sprtype[k] = pmx.SPRING_NONLINEAR_ELASTIC
    m.spridx[S], m.dashidx[S], m.unidx[S] = cs, ds, us


def _not_here(name):
    def f(*a, **k):
        raise NotImplementedError(f"{name} is not on the MI355X step path")
    f.__name__ = name
    return f


ANALYTICAL = _not_here("ANALYTICAL")


def OUTPUT(entities=None, subset=None, mode=None, format=None):
    """Output entities, modes and formats for a subset of particles, or for
    the particles no subset lists (input.cpp:3560-3956).  Every call adds a
    record (with an empty subset when none is given), as the reference does.
    DUMP and VTK files are written at DEM's output points; XDMF and MED need
    libraries this build has not got (NotImplementedError)."""
    if entities is not None and not isinstance(entities, list):
        raise TypeError("'entities' must be a list")
    if subset is not None and not (isinstance(subset, list) or _is_number(subset)):
        raise TypeError("'subset' must be a list or a number")
    if mode is not None and not isinstance(mode, (str, list)):
        raise TypeError("'mode' must be a string or a list")
    if format is not None and not isinstance(format, (str, list)):
        raise TypeError("'format' must be a string or a list")
    parts = []
    if subset is not None:
        for k in (subset if isinstance(subset, list) else [subset]):
            k = int(k)
            if k < 0 or k >= len(_S.parts):
                raise ValueError("Invalid particle number")
            parts.append(k)
    rec_ent, rec_mode = 0, 0
    if mode is not None:
        items = [mode] if isinstance(mode, str) else mode
        if isinstance(mode, list):
            if subset is None:
                _S.outrest[1] = 0
        for it in items:
            if not isinstance(it, str):
                raise TypeError("mode list item is not a string")
            if it not in OUT_MODES:
                raise ValueError("Invalid mode")
            b = OUT_MODES[it]
            if subset is not None or it in ("SL", "ST", "JT"):
                # SL, ST and JT set the record's own mode even without a subset
                rec_mode = rec_mode | b if isinstance(mode, list) else b
            elif it in ("SPH", "MESH"):
                _S.outrest[1] = (_S.outrest[1] | b) if isinstance(mode, list) else b
            # RB and CD without a subset write outmode[1], not the default
            # modes (input.cpp:3652-3661, 3698-3707): no effect here
    else:
        if subset is not None:
            rec_mode = OUT_ALL_MODES
        else:
            _S.outrest[1] = OUT_ALL_MODES
    if entities is not None:
        if subset is None:
            _S.outrest[0] = 0
        for it in entities:
            if not isinstance(it, str) or it not in OUT_ENTITIES:
                raise ValueError("Invalid entity")
            if subset is not None:
                rec_ent |= OUT_ENTITIES[it]
            else:
                _S.outrest[0] |= OUT_ENTITIES[it]
    else:
        if subset is not None:
            rec_ent = 0xffffffff
        else:
            _S.outrest[0] = 0xffffffff
    if format is not None:
        names = {"DUMP": OUT_FORMAT_DUMP, "VTK": OUT_FORMAT_VTK, "XDMF": OUT_FORMAT_XDMF, "MED": OUT_FORMAT_MED}
        fmt = 0
        for it in ([format] if isinstance(format, str) else format):
            if it not in names:
                raise ValueError("Invalid format")
            fmt |= names[it]
        if fmt & (OUT_FORMAT_XDMF | OUT_FORMAT_MED):
            raise NotImplementedError("XDMF and MED output need HDF5 / MED libraries, not in this build; "
                                      "use format=['DUMP', 'VTK']")
        _S.outformat = fmt
    _S.outputs.append((parts, rec_ent, rec_mode))


def _output_path(prefix):
    """The reference's output path: the input deck's path without '.py'; a
    prefix replaces its file name part (input.cpp:4140-4153,
    parmec.cpp:2328-2351)."""
    import os
    import sys
    if _S.output_path is None:
        deck = sys.argv[0] if sys.argv and sys.argv[0].endswith(".py") else "parmec.py"
        _S.output_path = deck[:-3]
    if prefix is not None:
        head = os.path.dirname(_S.output_path)
        _S.output_path = os.path.join(head, prefix) if head else prefix
    return _S.output_path


def _write_output(m, t, prefix_path):
    """output_files at curtime t for the state m (downloaded)."""
    from . import output
    flags = np.full(m.parnum, _OUTREST, dtype=np.int32)
    for parts, _, _ in _S.outputs:
        for k in parts:
            flags[k] &= ~_OUTREST
    recs = _S.springs
    order = sorted(range(len(recs)), key=lambda i: (recs[i]["part"][0], recs[i]["part"][1], i))
    trs = _S.torsion
    torder = sorted(range(len(trs)), key=lambda i: (trs[i]["part"][0], trs[i]["part"][1], i))
    output.write(m, t, _S.output_frame, prefix_path, _S.outformat, _S.outputs, _S.outrest, flags, order,
                 (torder, [trs[i]["refpnt"] for i in torder]))
    _S.output_frame += 1


def _build_model():
    """SoA model of everything created so far (the buffers of parmec.h):
    particles in creation order (SPHERE and MESH share the numbering),
    spheres as ellipsoids, mesh nodes and faces, triangles of obstacles and
    meshes in creation order."""
    n, e, t = len(_S.parts), len(_S.sph), len(_S.tris)
    old = _S.model
    npairs = len(_S.pairs)
    m = pmx.Model(parnum=n, ellnum=e, trinum=t, pairnum=npairs)
    m.flags[:] = 2   # OUTREST (constants.h:39), as SPHERE and MESH set it
    if e:
        a = np.array([s_[:4] for s_ in _S.sph], dtype=np.float64)
        mat = np.array([s_[4] for s_ in _S.sph])
        col = np.array([s_[5] for s_ in _S.sph], dtype=np.int32)
        part = np.array([s_[6] for s_ in _S.sph], dtype=np.int32)
        dens = np.array([_S.materials[k][_DENSITY] for k in mat])
        r = a[:, 3]
        m.position[:3, part] = a[:, :3].T
        m.position[3:, part] = a[:, :3].T
        m.center[:3] = a[:, :3].T
        m.center[3:] = a[:, :3].T
        m.radii[0] = r
        m.radii[1:] = -1.0
        m.part[:] = part
        m.ellcol[:] = col
        volume = (4.0 / 3.0) * math.pi * r * r * r
        m.mass[part] = volume * dens
        m.invm[part] = 1.0 / m.mass[part]
        j = 0.4 * m.mass[part] * r * r
        for k in (0, 4, 8):
            m.inertia[k, part] = j
    nodnum = sum(len(ms["nodes"]) for ms in _S.meshes)
    facnum = sum(len(ms["triangles"]) for ms in _S.meshes)
    m.set_mesh(nodnum, facnum)
    no = fo = 0
    for ms in _S.meshes:
        p = ms["part"]
        m.position[:3, p] = ms["center"]
        m.position[3:, p] = ms["center"]
        m.mass[p] = ms["mass"]
        m.invm[p] = 1.0 / ms["mass"]
        m.inertia[:, p] = ms["inertia"]
        m.inverse[:, p] = np.linalg.inv(np.array(ms["inertia"]).reshape(3, 3)).reshape(-1)
        x = np.array(ms["nodes"], dtype=np.float64)
        k = len(x)
        m.nodes[:3, no:no + k] = x.T
        m.nodes[3:, no:no + k] = x.T
        m.nodpart[no:no + k] = p
        tri = np.array([abc for abc, _ in ms["triangles"]], dtype=np.int32).reshape(-1, 3)
        f = len(tri)
        m.facnod[:, fo:fo + f] = (tri + no).T
        m.factri[fo:fo + f] = ms["first_triangle"] + np.arange(f, dtype=np.int32)
        no += k
        fo += f
    if old is not None:   # particles created earlier keep their evolved state
        k, ke, kn = old.parnum, old.ellnum, old.nodnum
        for name in ("angular", "linear", "rotation", "position", "force", "torque"):
            getattr(m, name)[:, :k] = getattr(old, name)[:, :k]
        m.center[:, :ke] = old.center[:, :ke]
        m.nodes[:, :kn] = old.nodes[:, :kn]
        m.curtime, m.curstep, m.stepnum = old.curtime, old.curstep, old.stepnum
    for p, (lin, ang) in _S.vel.items():
        if lin:
            m.linear[:, p] = lin
        if ang:
            m.angular[3:, p] = ang
            L = m.rotation[:, p]
            o = ang
            m.angular[0, p] = L[0] * o[0] + L[1] * o[1] + L[2] * o[2]
            m.angular[1, p] = L[3] * o[0] + L[4] * o[1] + L[5] * o[2]
            m.angular[2, p] = L[6] * o[0] + L[7] * o[1] + L[8] * o[2]
    _S.vel.clear()
    if t:
        m.tri[:] = np.array(_S.tris, dtype=np.float64).T
        m.tricol[:] = np.array(_S.tricol, dtype=np.int32)
        m.triobs[:] = np.array(_S.triobs, dtype=np.int32)
    m.set_prescribed(len(_S.prescribed))
    cols = {}
    for k, (p, lin, ang, lk, ak) in enumerate(_S.prescribed):
        m.prspart[k] = p
        m.prskind[0, k] = lk if lin is not None else -1
        m.prskind[1, k] = ak if ang is not None else -1
        m.prscol[k] = cols.setdefault((id(lin), id(ang)), len(cols))   # entries sharing histories share a column
    m.set_obstacles(len(_S.obstacles))
    for k, (a, b, pnt, lin, ang) in enumerate(_S.obstacles):
        m.trirng[k] = (a, b)
        m.obspnt[k] = pnt
    if old is not None:   # moved obstacle triangles and obstacles keep their state
        m.tri[:, :old.trinum] = old.tri
        k = old.obsnum
        m.obspnt[:k], m.obslin[:k], m.obsang[:k] = old.obspnt, old.obslin, old.obsang
    if _S.springs:
        _spring_arrays(m)
        if old is not None and old.sprnum == m.sprnum:
            # springs keep the state UNSPRING gave them; a spring an UNSPRING
            # made since the last run is controlled by it (-2)
            m.unspring[:] = np.where((m.unspring == -2) & (old.unspring == -3), -2, old.unspring)
    _f3_arrays(m)
    # sort_materials, parmec.cpp:1636-1663: pairs ordered by colours
    pairs = sorted(_S.pairs, key=lambda p: (p[0], p[1]))
    for i, (c1, c2, par) in enumerate(pairs):
        m.pairs[2 * i], m.pairs[2 * i + 1] = c1, c2
        m.ikind[i] = pmx.GRANULAR_FORCE
        m.iparam[:, i] = par
    return m


def _history(x, name):
    """A PRESCRIBE history: a callable of time or a tuple of 3 TSERIES numbers."""
    if x is None:
        return None
    if callable(x):
        return x
    if isinstance(x, tuple):
        if len(x) != 3:
            raise ValueError(f"'{name}' tuple size != 3")
        ids = [int(v) for v in x]
        if any(i < 0 or i >= len(_S.tms) for i in ids):
            raise ValueError(f"'{name}' time series number out of range")
        series = [_S.tms[i] for i in ids]
        return lambda t: tuple(s(t) for s in series)
    raise ValueError(f"'{name}' is neither a tuple nor a callback")


def PRESCRIBE(parnum, linear=None, angular=None, kind="vv"):
    """Prescribe the linear and angular velocity ('v') or acceleration ('a')
    of a particle (input.cpp:2647-2795); kind is 'vv', 'va', 'av' or 'aa'
    (linear, angular).  A history is a callable of time returning a 3-tuple,
    or a tuple of 3 TSERIES numbers."""
    n = len(_S.parts)
    if not isinstance(parnum, int) or parnum < 0 or parnum >= n:
        raise ValueError("'parnum' must be >= 0 and < the number of particles")
    if not isinstance(kind, str):
        raise TypeError("'kind' must be a string")
    if kind not in ("vv", "va", "av", "aa"):
        raise ValueError("Invalid time history kind")
    lin, ang = _history(linear, "linear"), _history(angular, "angular")
    if lin is None and ang is None:
        return
    _S.prescribed.append((parnum, lin, ang, 0 if kind[0] == "v" else 1, 0 if kind[1] == "v" else 1))
    _S.dirty = True


_SPRING_ENTITIES = pmx.HIS_ENTITIES[pmx.HIS_SPRING_FIRST:pmx.HIS_TORSION_FIRST]
_TORSION_ENTITIES = pmx.HIS_ENTITIES[pmx.HIS_TORSION_FIRST:pmx.HIS_JOINT_FIRST]
_JOINT_ENTITIES = pmx.HIS_ENTITIES[pmx.HIS_JOINT_FIRST:]


def HISTORY(entity, source=None, point=None, h5file=None, h5last=False):
    """Time history of an entity; returns a list that DEM appends to at each
    history interval (input.cpp:3157-3555, output.cpp:5143-5520).  Entities:
    TIME; PX PY PZ |P| DX DY DZ |D| VX VY VZ |V| OX OY OZ |O| FX FY FZ |F| TX
    TY TZ |T| over a particle number or list (averaged), or one particle's
    body point (point = (x, y, z), referential); LENGTH STROKE F SF FF SS
    over spring numbers; ZDIR_X .. XDIR_Z, ROLL PITCH YAW, TRQTOT_R/P/Y and
    TRQSPR_R/P/Y over torsion spring numbers; JREAC_X/Y/Z and |JREAC| over
    joint numbers.  SS reads the spring's UNSPRING state and ZDIR/XDIR the
    torsion spring's current axes (the reference indexes both through the
    torsion spring map, output.cpp:5400-5440).  Sphere and box sources and
    .h5 post-processing are not on this path."""
    if not isinstance(entity, str):
        raise TypeError("'entity' must be a string")
    if point is not None and not (isinstance(point, tuple) and len(point) == 3):
        raise ValueError("'point' must be a tuple of size 3")
    if h5file is not None:
        raise NotImplementedError("HISTORY from .h5 files is not on the MI355X step path")
    if entity not in pmx.HIS_ENTITIES:
        raise ValueError("Invalid entity")
    torsion, joint = entity in _TORSION_ENTITIES, entity in _JOINT_ENTITIES
    spring = entity in _SPRING_ENTITIES or torsion or joint
    nmax = (len(_S.torsion) if torsion else len(_S.joints) if joint else len(_S.springs)) if spring else len(_S.parts)
    if source is None:                       # useful when entity is 'TIME'
        if not _S.parts:
            raise ValueError("No particle has been defined")
        lst = [0]
    elif isinstance(source, int) and not isinstance(source, bool):
        lst = [source]
    elif isinstance(source, list):
        lst = [int(k) for k in source]
    elif isinstance(source, tuple):
        if spring:
            raise ValueError("Invalid source kind for a spring/joint entity")
        if len(source) not in (4, 6):
            raise ValueError("Invalid source tuple size")
        raise NotImplementedError(("Sphere" if len(source) == 4 else "Box") +
                                  " based time history is not yet implemented")
    else:
        raise TypeError("'source' must be a number, a list or a tuple")
    for k in lst:
        if k < 0 or k >= nmax:
            raise ValueError(("Torsion spring" if torsion else "Joint" if joint else "Spring") + " index out of range"
                             if spring else "Particle index out of range")
    pt = tuple(float(x) for x in point) if point is not None and len(lst) == 1 and not spring else None
    out = []
    _S.histories.append((entity, lst, pt, out))
    return out


def _history_list(entity, lst):
    """A history's source numbers as the model arrays index them: springs
    and torsion springs by their place in the sorted arrays (sprmap,
    trqsprmap), particles and joints by number."""
    if entity in _SPRING_ENTITIES:
        return [_S.sprmap[k] for k in lst]
    if entity in _TORSION_ENTITIES:
        return [_S.trqsprmap[k] for k in lst]
    return lst


def _interval(x, step):
    """One DEM output interval: a number, a callable of time or a TSERIES
    number (input.cpp:3977-4080); None: the step."""
    if x is None:
        return lambda t: step
    if callable(x):
        return lambda t: float(x(t))
    if isinstance(x, int) and not isinstance(x, bool):
        if x < 0 or x >= len(_S.tms):
            raise ValueError("Invalid output interval TSERIES number")
        ts = _S.tms[x]
        return lambda t: float(ts(t))
    if _is_number(x):
        if x < 0.0:
            raise ValueError("Invalid, negative, output interval")
        return lambda t: float(x)
    raise ValueError("Invalid output interval")


def _history_rows(times, curtime_history, dt_his):
    """The steps whose end state dem() samples (parmec.cpp:2509-2514): at
    step k, with curtime = times[k], a sample is taken when curtime >=
    curtime_history + interval(curtime), and curtime_history then advances by
    that interval.  Returns (row of each step or -1, sampled curtimes, the new
    curtime_history)."""
    rows = np.full(max(len(times), 1), -1, dtype=np.int32)
    tsamp = []
    ch = curtime_history
    for k, t in enumerate(times):
        d = dt_his(t)
        if t >= ch + d:
            rows[k] = len(tsamp)
            tsamp.append(t)
            ch += d
    return rows, tsamp, ch


def _engine_model():
    """The engine holding the deck's model, uploading the model when the
    deck changed since the last upload."""
    if _S.engine is None:
        _S.engine = pmx.Engine(device=_S.options["device"], row_capacity=_S.options["row_capacity"], critical=1)
        _S.dirty = True
    if _S.dirty or _S.model is None:
        _S.model = _build_model()
        _S.engine.upload(_S.model)
        _S.dirty = False
    return _S.engine, _S.model


def _inputs_at():
    """Evaluators of dem()'s host inputs at a time (input.cpp:4199-4503):
    gravity, damping, obstacle velocities (angular, linear) and the
    prescribed values of each history column; None where the deck has
    nothing time dependent of that kind."""
    m = _S.model
    grav = None
    if any(f is not None for f in _S.gravfunc):
        def grav(t):
            return [float(_S.gravfunc[k](t)) if _S.gravfunc[k] is not None else _S.gravity[k] for k in range(3)]
    damp = None
    if _S.lindamp is not None or _S.angdamp is not None:
        def damp(t):
            d = np.zeros(6)
            if _S.lindamp is not None:
                d[:3] = _S.lindamp(t)
            if _S.angdamp is not None:
                d[3:] = _S.angdamp(t)
            return d
    vel = None
    if _S.obstacles:
        def vel(t):
            row = np.zeros((len(_S.obstacles), 6))
            for k, (_, _, _, lin, ang) in enumerate(_S.obstacles):
                if ang is not None:
                    row[k, :3] = [float(x) for x in ang(t)]
                if lin is not None:
                    row[k, 3:] = [float(x) for x in lin(t)]
            return row
    prs = None
    if _S.prescribed:
        ncol = int(m.prscol.max()) + 1
        cols = {}
        for k, (_, lin, ang, _, _) in enumerate(_S.prescribed):
            cols.setdefault(int(m.prscol[k]), (lin, ang))

        def prs(t):
            row = np.zeros((ncol, 6))
            for c, (lin, ang) in cols.items():
                if lin is not None:
                    row[c, :3] = [float(v) for v in lin(t)]
                if ang is not None:
                    row[c, 3:] = [float(v) for v in ang(t)]
            return row
    return grav, damp, vel, prs


def _intervals(interval, step):
    """(output, history) interval evaluators of DEM's interval argument."""
    if isinstance(interval, tuple):
        if len(interval) != 2:
            raise ValueError("Invalid output interval")
        return _interval(interval[0], step), _interval(interval[1], step)
    return _interval(interval, step), _interval(interval, step)


def DEM(duration, step, interval=None, prefix=None, adaptive=0.0):
    """Run the DEM time stepping; returns the wall time in seconds
    (input.cpp:3958-4091, parmec.cpp:2315-2530).  interval = (output, history)
    or one value for both (number, callable or TSERIES number; default the
    step): HISTORY lists are appended at each history interval, and the
    DUMP/VTK files OUTPUT asked for are written at each output interval
    (under the deck's path without '.py', or the prefix).  adaptive in
    (0, 1] takes each step as that fraction of the step's critical estimate
    (parmec.cpp:2444-2447).""" 
    if not duration > 0:
        raise ValueError("'duration' must be positive")
    if not step > 0:
        raise ValueError("'step' must be positive")
    if prefix is not None and not isinstance(prefix, str):
        raise TypeError("'prefix' must be a string")
    if not (0.0 <= adaptive <= 1.0):
        raise ValueError("'adaptive' must be >= 0 and <= 1")
    engine, m = _engine_model()
    h = step if m.curtime == 0.0 else m.curstep
    grav, damp, vel, prs = _inputs_at()
    dt_out, dt_his = _intervals(interval, step)
    obstab0 = vel(0.5 * h) if vel is not None and m.curtime == 0.0 else None
    if _S.histories:
        ents = [e for e, _, _, _ in _S.histories]
        engine.set_history(ents, [_history_list(e, l) for e, l, _, _ in _S.histories],
                           [p for _, _, p, _ in _S.histories])
        if m.curtime == 0.0:               # the sample before the first step (parmec.cpp:2386-2394)
            v = engine.history_sample()
            for i, (e, _, _, out) in enumerate(_S.histories):
                out.append(0.0 if e == "TIME" else float(v[i]))
    files = bool(_S.outformat & (OUT_FORMAT_DUMP | OUT_FORMAT_VTK))
    path = _output_path(prefix) if files else None
    if files and m.curtime == 0.0:         # output_files before the first step (parmec.cpp:2386-2392)
        _write_output(m, 0.0, path)
    if adaptive > 0.0:
        return _dem_adaptive(engine, m, duration, step, adaptive, grav, damp, vel, prs, obstab0, dt_out, dt_his,
                             path)
    times = pmx.step_times(duration, h, m.curtime)
    n = len(times)
    gtab = np.array([grav(t) for t in times]).reshape(n, 3) if grav is not None else None
    dtab = np.array([damp(t) for t in times]).reshape(n, 6) if damp is not None else None
    # obstaclev at curtime+h of each step, prescribed values at curtime
    obstab = np.array([vel(t + h) for t in times]).reshape(n, len(_S.obstacles), 6) if vel is not None else None
    prstab = np.array([prs(t) for t in times]).reshape(n, -1, 6) if prs is not None else None
    hisrow = hisout = None
    if _S.histories:
        # which steps' end states are sampled (parmec.cpp:2509-2514): the
        # history interval at each step's curtime, crossed by curtime
        hisrow, tsamp, _S.curtime_history = _history_rows(times, _S.curtime_history, dt_his)
        hisout = np.zeros((max(len(tsamp), 1), len(_S.histories)))
        if not tsamp:
            hisrow = hisout = None
    outsteps, on_output = None, None
    if files:
        # output_files at the end state of every step whose curtime reaches
        # curtime_output + interval (parmec.cpp:2502-2507)
        orow, tout, _S.curtime_output = _history_rows(times, _S.curtime_output, dt_out)
        outsteps = [k for k in range(n) if orow[k] >= 0]

        def on_output(k):
            snap = m.copy()
            engine.download(snap)
            _write_output(snap, float(times[k]), path)
    f = pmx.forcing(n, gravity=tuple(_S.gravity), gtab=gtab, dtab=dtab, obstab=obstab, obstab0=obstab0,
                    prstab=prstab, hisrow=hisrow, hisout=hisout, outsteps=outsteps, on_output=on_output)
    elapsed = engine.dem(float(duration), float(step), f)
    engine.download(m)
    if hisrow is not None:
        for r, t in enumerate(tsamp):
            for i, (e, _, _, out) in enumerate(_S.histories):
                out.append(float(t) if e == "TIME" else float(hisout[r, i]))
    return elapsed


def _dem_adaptive(engine, m, duration, step, adaptive, grav, damp, vel, prs, obstab0, dt_out, dt_his, path):
    """DEM with the adaptive step: the inputs of each step are evaluated at
    its curtime as the step is reached, and the history samples and output
    files are taken at pauses after the steps that reach an interval (the
    same states the tabulated path samples inside the step loop)."""
    due = {}

    def on_step(k, t, h0, inp):
        if grav is not None:
            inp.gravity[:] = grav(t)
        if damp is not None:
            inp.damping[:] = damp(t)
        if vel is not None:
            inp.obsvel[:] = vel(t + h0)
        if prs is not None:
            inp.prs[:] = prs(t)
        files = his = False
        if path is not None:
            d = dt_out(t)
            if t >= _S.curtime_output + d:
                _S.curtime_output += d
                files = True
        if _S.histories:
            d = dt_his(t)
            if t >= _S.curtime_history + d:
                _S.curtime_history += d
                his = True
        if files or his:
            due[k] = (t, files, his)
            inp.output = 1

    def on_output(k):
        t, files, his = due.pop(k)
        if files:
            snap = m.copy()
            engine.download(snap)
            _write_output(snap, t, path)
        if his:
            v = engine.history_sample()
            for i, (e, _, _, out) in enumerate(_S.histories):
                out.append(t if e == "TIME" else float(v[i]))

    obsnum = len(_S.obstacles)
    prscols = int(m.prscol.max()) + 1 if _S.prescribed else 0
    f = pmx.forcing(0, gravity=tuple(_S.gravity), obstab0=obstab0, on_step=on_step, on_output=on_output,
                    obsnum=obsnum, prscols=prscols)
    elapsed = engine.dem_adaptive(float(duration), float(step), float(adaptive), f)
    engine.download(m)
    return elapsed


def CRITICAL(perspring=-1, perparticle=-1):
    """Critical time step estimates (input.cpp:3034-3154): with no argument
    the smallest per-particle estimate; perspring = n and/or perparticle = n
    give the n smallest as lists of (step, index, omega, ratio), sorted by
    step (a pair of lists when both are asked for).  The per-particle
    estimate takes the contacts of the last step taken, so before the first
    DEM only springs load a particle, and a particle nothing loads estimates
    DBL_MAX.  NaN estimates (a loaded particle with zero stiffness) sort
    last, and equal ones by index."""
    if not isinstance(perspring, int) or not isinstance(perparticle, int):
        raise TypeError("'perspring' and 'perparticle' must be integers")
    engine, m = _engine_model()
    if perspring < -1 or perspring > m.sprnum:
        raise ValueError("Parameter 'perspring' is out of range")
    if perparticle < -1 or perparticle > m.parnum:
        raise ValueError("Parameter 'perparticle' is out of range")

    def ranked(which, count):
        h, o, r = engine.critical(which)
        order = sorted(range(len(h)), key=lambda i: (h[i] != h[i], h[i] if h[i] == h[i] else 0.0, i))
        return [(float(h[i]), int(i), float(o[i]), float(r[i])) for i in order[:count]]

    sp = ranked(pmx.CRIT_SPRING, perspring) if perspring >= 0 else None
    pp = ranked(pmx.CRIT_PARTICLE, perparticle) if perparticle >= 0 else None
    if sp is not None and pp is not None:
        return (sp, pp)
    if sp is not None:
        return sp
    if pp is not None:
        return pp
    v = ranked(pmx.CRIT_PARTICLE, 1)
    return v[0][0] if v else 1.7976931348623157e308


def _use_engine(engine):
    """Run DEM on a caller-supplied engine object (same methods as pmx.Engine)."""
    _S.engine = engine
    _S.dirty = True


def state():
    """The SoA model (numpy arrays shaped like parmec.h's buffers)."""
    if _S.model is None:
        _S.model = _build_model()
    return _S.model


def contacts():
    """Current contact pairs (master ellipsoid, slave ellipsoid or -(triangle+1), state)."""
    return _S.engine.contacts() if _S.engine is not None else None
