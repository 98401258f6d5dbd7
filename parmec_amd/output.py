"""Output files of DEM() at its output points (output.cpp:3968-5142):
LAMMPS-style DUMP files of the spheres and legacy ASCII VTK files of the
triangles (meshes and obstacles), rigid bodies and linear springs, from a
snapshot of the resident state.

The file names and the contents follow the reference writers:
  <path><k>.dump            spheres, appended frame after frame (SPH mode)
  <path><k>.vtk.<frame>     triangles (MESH mode)
  <path><k>rb.vtk.<frame>   rigid bodies: particle positions (RB mode)
  <path><k>sl.vtk.<frame>   linear springs (SL mode)
  <path><k>st.vtk.<frame>   torsion springs (ST mode)
  <path><k>jt.vtk.<frame>   ball joints (JT mode)
where k = 0 is the set of particles no OUTPUT subset lists and k = j + 1 the
j-th OUTPUT record.  Numbers are written as C++ streams write doubles
(six significant digits, like %g).  XDMF/HDF5 and MED are not written (no
HDF5 or MED library in this build).
"""
import numpy as np

from . import pmx

E = {"NUMBER": 1, "COLOR": 2, "DISPL": 4, "ORIENT": 8, "ORIENT1": 16, "ORIENT2": 32, "ORIENT3": 64, "LINVEL": 128,
     "ANGVEL": 256, "FORCE": 512, "TORQUE": 1024, "F": 2048, "SF": 16384, "LENGTH": 131072, "SS": 262144,
     "FF": 524288, "XDIR": 1048576, "YDIR": 2097152, "ZDIR": 4194304, "TRQROT": 8388608, "TRQTOT": 16777216,
     "TRQSPR": 33554432, "JREAC": 67108864}
M_SPH, M_MESH, M_RB, M_SL, M_ST, M_JT = 1, 2, 4, 16, 32, 64
_REAL_MAX = 1.7976931348623157e308
OUTREST = 2


def g(x):
    """A double as `ostream << x` writes it (precision 6, %g)."""
    return "%g" % x


def vec(a, b, c):
    return f"{g(a)} {g(b)} {g(c)}"


def _dump_dataset(m, t, sel, ent, out):
    """output_dump_dataset, output.cpp:54-143 (sel: ellipsoid numbers)."""
    out.write(f"ITEM: TIMESTEP\n{g(t)}\nITEM: NUMBER OF ATOMS\n{len(sel)}\nITEM: BOX BOUNDS\n0 1\n0 1\n0 1\n")
    out.write("ITEM: ATOMS id x y z radius")
    if ent & E["LINVEL"]:
        out.write(" vx vy vz")
    if ent & E["ANGVEL"]:
        out.write(" omegax omegay omegaz")
    if ent & E["FORCE"]:
        out.write(" fx fy fz")
    if ent & E["TORQUE"]:
        out.write(" tqx tqy tqz")
    out.write("\n")
    for j in sel:
        k = m.part[j]
        line = f"{j + 1} {vec(*m.center[:3, j])} {g(m.radii[0, j])}"
        if ent & E["LINVEL"]:
            line += " " + vec(*m.linear[:, k])
        if ent & E["ANGVEL"]:
            line += " " + vec(*m.angular[:3, k])
        if ent & E["FORCE"]:
            line += " " + vec(*m.force[:, k])
        if ent & E["TORQUE"]:
            line += " " + vec(*m.torque[:, k])
        out.write(line + "\n")


def _tri_point_vectors(m, tset, which):
    """Per-vertex displacement ('displ') or velocity ('linvel') of triangles
    (output.cpp:1741-1902): mesh particles from their motion, static
    obstacles zero, moving obstacles about their point."""
    lines = []
    for t in tset:
        p = [m.tri[3 * v:3 * v + 3, t] for v in range(3)]
        j = m.triobs[t]
        for q in p:
            if j >= 0:
                x = m.position[:3, j]
                if which == "displ":
                    L = m.rotation[:, j]
                    d = q - x
                    X = m.position[3:, j]
                    P = np.array([X[0] + (L[0] * d[0] + L[1] * d[1] + L[2] * d[2]),
                                  X[1] + (L[3] * d[0] + L[4] * d[1] + L[5] * d[2]),
                                  X[2] + (L[6] * d[0] + L[7] * d[1] + L[8] * d[2])])
                    lines.append(vec(*(q - P)))
                else:
                    o, v, a = m.angular[3:, j], m.linear[:, j], q - x
                    lines.append(vec(v[0] + o[1] * a[2] - o[2] * a[1], v[1] + o[2] * a[0] - o[0] * a[2],
                                     v[2] + o[0] * a[1] - o[1] * a[0]))
            elif j == -1:
                lines.append("0.0 0.0 0.0")
            else:
                k = -j - 2
                x = m.obspnt[k]
                if which == "displ":
                    lines.append(vec(*(q - x)))
                else:
                    o, v, a = m.obsang[k], m.obslin[k], q - x
                    lines.append(vec(v[0] + o[1] * a[2] - o[2] * a[1], v[1] + o[2] * a[0] - o[0] * a[2],
                                     v[2] + o[0] * a[1] - o[1] * a[0]))
    return lines


def _triangle_dataset(m, tset, ent, out):
    """vtk_triangle_dataset, output.cpp:1709-1977."""
    n = len(tset)
    out.write(f"DATASET UNSTRUCTURED_GRID\nPOINTS {3 * n} float\n")
    for t in tset:
        out.write(" ".join(g(m.tri[k, t]) for k in range(9)) + "\n")
    out.write(f"CELLS {n} {4 * n}\n")
    for i in range(n):
        out.write(f"3 {3 * i} {3 * i + 1} {3 * i + 2}\n")
    out.write(f"CELL_TYPES {n}\n" + "5\n" * n)
    if ent & (E["DISPL"] | E["LINVEL"]):
        out.write(f"POINT_DATA {3 * n}\n")
    if ent & E["DISPL"]:
        out.write("VECTORS displ float\n" + "".join(x + "\n" for x in _tri_point_vectors(m, tset, "displ")))
    if ent & E["LINVEL"]:
        out.write("VECTORS linvel float\n" + "".join(x + "\n" for x in _tri_point_vectors(m, tset, "linvel")))
    if ent & (E["NUMBER"] | E["COLOR"] | E["ANGVEL"] | E["FORCE"] | E["TORQUE"]):
        out.write(f"CELL_DATA {n}\n")
    if ent & E["NUMBER"]:
        out.write("SCALARS number int\nLOOKUP_TABLE default\n" + "".join(f"{m.triobs[t]}\n" for t in tset))
    if ent & E["COLOR"]:
        out.write("SCALARS colors int\nLOOKUP_TABLE default\n" + "".join(f"{m.tricol[t]}\n" for t in tset))

    def per_owner(name, particle, obstacle):
        out.write(f"VECTORS {name} float\n")
        for t in tset:
            j = m.triobs[t]
            if j >= 0:
                out.write(vec(*particle(j)) + "\n")
            elif j == -1 or obstacle is None:
                out.write("0.0 0.0 0.0\n")
            else:
                out.write(vec(*obstacle(-j - 2)) + "\n")
    if ent & E["ANGVEL"]:
        per_owner("angvel", lambda j: m.angular[3:, j], lambda k: m.obsang[k])
    if ent & E["FORCE"]:
        per_owner("force", lambda j: m.force[:, j], None)
    if ent & E["TORQUE"]:
        per_owner("torque", lambda j: m.torque[:, j], None)


def _rb_dataset(m, pset, ent, out):
    """output_rb_dataset, output.cpp:146-267 (particle numbers)."""
    n = len(pset)
    out.write(f"DATASET UNSTRUCTURED_GRID\nPOINTS {n} float\n")
    for j in pset:
        out.write(vec(*m.position[:3, j]) + "\n")
    if ent & (E["NUMBER"] | E["DISPL"] | E["ORIENT1"] | E["ORIENT2"] | E["ORIENT3"] | E["LINVEL"] | E["ANGVEL"] |
              E["FORCE"] | E["TORQUE"]):
        out.write(f"POINT_DATA {n}\n")
    if ent & E["NUMBER"]:
        out.write("SCALARS NUMBER int\nLOOKUP_TABLE default\n" + "".join(f"{j}\n" for j in pset))
    rows = (("DISPL", "DISPL", lambda j: m.position[:3, j] - m.position[3:, j]),
            ("ORIENT1", "ORIENT1", lambda j: m.rotation[0:3, j]),
            ("ORIENT2", "ORIENT2", lambda j: m.rotation[3:6, j]),
            ("ORIENT3", "ORIENT3", lambda j: m.rotation[6:9, j]),
            ("LINVEL", "LINVEL", lambda j: m.linear[:, j]),
            ("ANGVEL", "ANGVEL", lambda j: m.angular[:3, j]),
            ("FORCE", "FORCE", lambda j: m.force[:, j]),
            ("TORQUE", "TORQUE", lambda j: m.torque[:, j]))
    for key, name, f in rows:
        if ent & E[key]:
            out.write(f"VECTORS {name} float\n" + "".join(vec(*f(j)) + "\n" for j in pset))


def _spring_dataset(m, sset, sprid, ent, out):
    """vtk_linear_spring_dataset, output.cpp:2855-2966 (spring numbers in
    the sorted arrays; sprid gives each one's creation number)."""
    n = len(sset)
    out.write(f"DATASET UNSTRUCTURED_GRID\nPOINTS {n} float\n")
    for j in sset:
        out.write(vec(*(0.5 * (m.sprpnt[0, :3, j] + m.sprpnt[1, :3, j]))) + "\n")
    if ent & (E["NUMBER"] | E["DISPL"] | E["F"] | E["SF"] | 65536 | E["SS"] | E["FF"]):
        out.write(f"POINT_DATA {n}\n")

    def scal(name, f):
        out.write(f"SCALARS {name} float\nLOOKUP_TABLE default\n" + "".join(g(f(j)) + "\n" for j in sset))
    if ent & E["NUMBER"]:
        out.write("SCALARS number int\nLOOKUP_TABLE default\n" + "".join(f"{sprid[j]}\n" for j in sset))
    if ent & E["DISPL"]:
        scal("displ", lambda j: m.stroke[0, j])
    if ent & E["LENGTH"]:
        scal("length", lambda j: float(np.sqrt(((m.sprpnt[1, :3, j] - m.sprpnt[0, :3, j]) ** 2).sum())))
    if ent & E["ORIENT"]:
        out.write("VECTORS orient float\n" + "".join(vec(*m.sprdir[:3, j]) + "\n" for j in sset))
    if ent & E["F"]:
        scal("F", lambda j: m.sprfrc[0, j])
    if ent & E["SF"]:
        scal("SF", lambda j: m.sprfrc[1, j])
    if ent & E["FF"]:
        scal("FF", lambda j: m.sprfrc[2, j])
    if ent & E["SS"]:
        scal("SS", lambda j: float(np.float32(m.unspring[j])))


def _body_point(m, k, P):
    """NVADDMUL (x, L, P - X): a referential point of particle k now."""
    d = [P[0] - m.position[3, k], P[1] - m.position[4, k], P[2] - m.position[5, k]]
    L = m.rotation[:, k]
    return [m.position[c, k] + (L[c] * d[0] + L[3 + c] * d[1] + L[6 + c] * d[2]) for c in range(3)]


def _torsion_dataset(m, tset, trqsprid, refpnt, ent, out):
    """vtk_torsional_spring_dataset, output.cpp:2969-3142 (torsion spring
    numbers in the sorted arrays; trqsprid their creation numbers, refpnt
    their reference points or REAL_MAX)."""
    n = len(tset)
    out.write(f"DATASET UNSTRUCTURED_GRID\nPOINTS {n} float\n")
    for j in tset:
        k, l = m.trqsprpart[0, j], m.trqsprpart[1, j]
        r = refpnt[j]
        if r[0] != _REAL_MAX:
            p = _body_point(m, k, r)
        elif l == -1:
            p = list(m.position[:3, k])
        else:
            p = [0.5 * (m.position[c, k] + m.position[c, l]) for c in range(3)]
        out.write(vec(*p) + "\n")
    if ent & (E["NUMBER"] | E["ZDIR"] | E["XDIR"] | E["YDIR"] | E["TRQROT"] | E["TRQTOT"] | E["TRQSPR"]):
        out.write(f"POINT_DATA {n}\n")
    if ent & E["NUMBER"]:
        out.write("SCALARS number int\nLOOKUP_TABLE default\n" + "".join(f"{trqsprid[j]}\n" for j in tset))

    def axes(j):
        z, x = m.trqzdir1[:, j], m.trqxdir1[:, j]
        y = [z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]]
        return x, y, z

    def frame(j, a):
        x, y, z = axes(j)
        return [x[c] * a[0, j] + y[c] * a[1, j] + z[c] * a[2, j] for c in range(3)]
    if ent & E["ZDIR"]:
        out.write("VECTORS ZDIR float\n" + "".join(vec(*m.trqzdir1[:, j]) + "\n" for j in tset))
    if ent & E["XDIR"]:
        out.write("VECTORS XDIR float\n" + "".join(vec(*m.trqxdir1[:, j]) + "\n" for j in tset))
    if ent & E["YDIR"]:
        out.write("VECTORS YDIR float\n" + "".join(vec(*axes(j)[1]) + "\n" for j in tset))
    if ent & E["TRQROT"]:
        out.write("VECTORS TRQROT float\n" + "".join(vec(*frame(j, m.trqrpy)) + "\n" for j in tset))
    if ent & E["TRQTOT"]:
        out.write("VECTORS TRQTOT float\n" + "".join(vec(*frame(j, m.trqrpytot)) + "\n" for j in tset))
    if ent & E["TRQSPR"]:
        out.write("VECTORS TRQSPR float\n" + "".join(vec(*frame(j, m.trqrpyspr)) + "\n" for j in tset))


def _joint_dataset(m, jset, ent, out):
    """vtk_joints_dataset, output.cpp:3145-3197: the joint point on part1
    now (the reference indexes part1 through the torsion spring array,
    output.cpp:3157; here the joint's own part1), its number and reaction."""
    n = len(jset)
    out.write(f"DATASET UNSTRUCTURED_GRID\nPOINTS {n} float\n")
    for j in jset:
        out.write(vec(*_body_point(m, m.jpart[0, j], m.jpoint[:, j])) + "\n")
    if ent & (E["NUMBER"] | E["JREAC"]):
        out.write(f"POINT_DATA {n}\n")
    if ent & E["NUMBER"]:
        out.write("SCALARS number int\nLOOKUP_TABLE default\n" + "".join(f"{j}\n" for j in jset))
    if ent & E["JREAC"]:
        out.write("VECTORS JREAC float\n" + "".join(vec(*m.jreac[:, j]) + "\n" for j in jset))


def _attached(parts_of, count, parts, flags):
    """The items (springs, torsion springs, joints) output for a record:
    those on its particles in item order per particle, or, for the rest,
    those whose part1 has any flag or whose part2 is unlisted
    (output.cpp:4551-4552, 4623-4624, 4695-4696)."""
    p1, p2 = parts_of
    if parts is None:
        return [k for k in range(count) if flags[p1[k]] != 0 or (p2[k] >= 0 and flags[p2[k]] & OUTREST)]
    return [k for p in parts for k in range(count) if p1[k] == p or p2[k] == p]


def write(m, t, frame, path, outformat, records, outrest, flags, sprid, trq=None):
    """output_files (output.cpp:5127-5141) for the state m at curtime t:
    records = [(particles, entities, modes)] of the OUTPUT calls; outrest =
    (entities, modes) of the particles no record lists; flags: particle
    flags with OUTREST cleared for listed particles; trq = (creation numbers,
    reference points) of the sorted torsion springs."""
    written = []
    sets = [(-1, None, outrest[0], outrest[1])] + [(j, parts, ent, mode) for j, (parts, ent, mode) in
                                                    enumerate(records)]
    if outformat & 1 and m.ellnum:                       # DUMP (output_dump_files)
        for j, parts, ent, mode in sets:
            if not (mode & M_SPH):
                continue
            if parts is None:
                sel = [e for e in range(m.ellnum) if flags[m.part[e]] & OUTREST]
            else:
                sel = [e for p in parts for e in range(m.ellnum) if m.part[e] == p]
            name = f"{path}{j + 1}.dump"
            with open(name, "a" if frame else "w") as out:
                _dump_dataset(m, t, sel, ent, out)
            written.append(name)
    if outformat & 2:                                    # VTK (output_vtk_files)
        if m.trinum:
            for j, parts, ent, mode in sets:
                if not (mode & M_MESH):
                    continue
                if parts is None:
                    tset = [k for k in range(m.trinum) if m.triobs[k] < 0 or flags[m.triobs[k]] & OUTREST]
                else:
                    tset = [k for p in parts for k in range(m.trinum) if m.triobs[k] == p]
                name = f"{path}{j + 1}.vtk.{frame}"
                with open(name, "w") as out:
                    out.write(f"# vtk DataFile Version 2.0\nPARMEC triangles output at time {g(t)}\nASCII\n")
                    if parts is None or tset:
                        _triangle_dataset(m, tset, ent, out)
                written.append(name)
        if m.parnum:
            for j, parts, ent, mode in sets:
                if not (mode & M_RB):
                    continue
                pset = [k for k in range(m.parnum) if flags[k] & OUTREST] if parts is None else list(parts)
                name = f"{path}{j + 1}rb.vtk.{frame}"
                with open(name, "w") as out:
                    out.write(f"# vtk DataFile Version 2.0\nPARMEC rigid bodies output at time {g(t)}\nASCII\n")
                    _rb_dataset(m, pset, ent, out)
                written.append(name)
        if m.sprnum:
            for j, parts, ent, mode in sets:
                if not (mode & M_SL):
                    continue
                if parts is None:
                    # the reference tests `flags[part1] && OUTREST` (any flag),
                    # or the second particle's OUTREST bit (output.cpp:4551-4552)
                    sset = [k for k in range(m.sprnum) if flags[m.sprpart[0, k]] != 0 or
                            (m.sprpart[1, k] >= 0 and flags[m.sprpart[1, k]] & OUTREST)]
                else:
                    sset = [k for p in parts for k in range(m.sprnum) if m.sprpart[0, k] == p or
                            m.sprpart[1, k] == p]
                name = f"{path}{j + 1}sl.vtk.{frame}"
                with open(name, "w") as out:
                    out.write(f"# vtk DataFile Version 2.0\nPARMEC linear springs output at time {g(t)}\nASCII\n")
                    if parts is None or sset:
                        _spring_dataset(m, sset, sprid, ent, out)
                written.append(name)
        if getattr(m, "trqsprnum", 0):
            ids, refs = trq if trq is not None else (list(range(m.trqsprnum)), [(_REAL_MAX,) * 3] * m.trqsprnum)
            for j, parts, ent, mode in sets:
                if not (mode & M_ST):
                    continue
                tset = _attached((m.trqsprpart[0], m.trqsprpart[1]), m.trqsprnum, parts, flags)
                name = f"{path}{j + 1}st.vtk.{frame}"
                with open(name, "w") as out:
                    out.write(f"# vtk DataFile Version 2.0\nPARMEC torsional springs output at time {g(t)}\nASCII\n")
                    if parts is None or tset:
                        _torsion_dataset(m, tset, ids, refs, ent, out)
                written.append(name)
        if getattr(m, "jnum", 0):
            for j, parts, ent, mode in sets:
                if not (mode & M_JT):
                    continue
                jset = _attached((m.jpart[0], m.jpart[1]), m.jnum, parts, flags)
                name = f"{path}{j + 1}jt.vtk.{frame}"
                with open(name, "w") as out:
                    out.write(f"# vtk DataFile Version 2.0\nPARMEC joints output at time {g(t)}\nASCII\n")
                    if parts is None or jset:
                        _joint_dataset(m, jset, ent, out)
                written.append(name)
    return written
