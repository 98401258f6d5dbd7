"""OUTPUT and the DUMP/VTK writers at DEM's output points (input.cpp:3560-
3956, output.cpp:54-267, 1709-1977, 2855-2966, 4406-4620, 5070-5125,
parmec.cpp:2386-2392, 2502-2507), driven by the oracle engine."""
import glob
import os

import numpy as np
import pytest

from parmec_amd import parmec as P
from tests.test_deck_mesh import cuboid_nodes, flat


@pytest.fixture(autouse=True)
def fresh():
    P.RESET()
    yield
    P.RESET()


def scene():
    """Two spheres above a meshed cube on a static floor, a spring from a
    sphere to the cube."""
    mat = P.MATERIAL(1000.0, 1e6, 0.25)
    a = P.SPHERE((0.2, 0.5, 1.3), 0.1, mat, 1)
    b = P.MESH(flat(cuboid_nodes((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))), [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 2)
    c = P.SPHERE((0.8, 0.5, 1.4), 0.1, mat, 1)
    P.OBSTACLE([(-2.0, -2.0, -0.01, 3.0, -2.0, -0.01, -2.0, 3.0, -0.01)], 3)
    P.SPRING(a, (0.2, 0.5, 1.3), b, (0.2, 0.5, 1.0), [-1.0, -1e3, 1.0, 1e3], dashpot=0.1)
    P.GRANULAR(0, 0, 1e5, 0.2, 0.3)
    P.GRAVITY(0.0, 0.0, -9.81)
    P.VELOCITY(c, (0.0, 0.0, -1.0))
    return a, b, c


def run(engine, tmp, interval, steps, h=1e-4, **out):
    from oracle import orc
    scene()
    P.OUTPUT(**out)
    P._use_engine(engine if engine is not None else orc.Oracle(1))
    P.DEM((steps - 0.5) * h, h, interval, prefix=os.path.join(tmp, "run"))
    return P.state()


def frames(path):
    with open(path) as f:
        return f.read().count("ITEM: TIMESTEP")


def test_files_at_output_points(tmp_path):
    """Interval 10 h over 100 steps: frames at t = 0 and after every step
    whose start time reaches the last output time + interval (10 + 1)."""
    m = run(None, str(tmp_path), 10e-4, 100, format=["DUMP", "VTK"])
    t = np.array([k * 1e-4 for k in range(100)])
    out_t, last = [], 0.0
    for x in t:
        if x >= last + 10e-4:
            out_t.append(x)
            last += 10e-4
    n = 1 + len(out_t)
    assert frames(os.path.join(tmp_path, "run0.dump")) == n
    assert len(glob.glob(os.path.join(tmp_path, "run0.vtk.*"))) == n
    assert len(glob.glob(os.path.join(tmp_path, "run0rb.vtk.*"))) == n
    assert len(glob.glob(os.path.join(tmp_path, "run0sl.vtk.*"))) == n
    # the last frame: the spheres' final centres, the mesh's final triangles
    lines = open(os.path.join(tmp_path, "run0.dump")).read().split("ITEM: TIMESTEP\n")[-1].splitlines()
    assert float(lines[0]) == pytest.approx(out_t[-1])
    rows = [l.split() for l in lines[8:]]
    assert [int(r[0]) for r in rows] == [1, 2]
    got = np.array([[float(x) for x in r[1:5]] for r in rows])
    if out_t[-1] == t[-1]:
        want = np.vstack([m.center[:3].T.T, m.radii[0]]).T
        np.testing.assert_allclose(got, want, rtol=1e-5)
    vtk = open(os.path.join(tmp_path, f"run0.vtk.{n - 1}")).read()
    assert "POINTS 39 float" in vtk and "CELLS 13 52" in vtk and "SCALARS colors int" in vtk
    rb = open(os.path.join(tmp_path, f"run0rb.vtk.{n - 1}")).read()
    assert "POINTS 3 float" in rb and "VECTORS LINVEL float" in rb
    sl = open(os.path.join(tmp_path, f"run0sl.vtk.{n - 1}")).read()
    assert "POINTS 1 float" in sl and "SCALARS F float" in sl


def test_subset_record(tmp_path):
    """A record listing particle 2 in RB mode with two entities; the other
    particles keep the default output."""
    from oracle import orc
    scene()
    P.OUTPUT(format="VTK")
    P.OUTPUT(entities=["NUMBER", "LINVEL"], subset=[2], mode=["RB"])
    P._use_engine(orc.Oracle(1))
    P.DEM(19.5e-4, 1e-4, 10e-4, prefix=os.path.join(str(tmp_path), "sub"))
    rb1 = open(os.path.join(tmp_path, "sub2rb.vtk.1")).read()      # record 1 (the second OUTPUT call)
    assert "POINTS 1 float" in rb1 and "SCALARS NUMBER int" in rb1 and "VECTORS LINVEL" in rb1
    assert "FORCE" not in rb1
    rb0 = open(os.path.join(tmp_path, "sub0rb.vtk.1")).read()      # particles 0 and 1
    assert "POINTS 2 float" in rb0
    assert not glob.glob(os.path.join(tmp_path, "sub*.dump"))


def test_output_argument_errors():
    scene()
    with pytest.raises(ValueError):
        P.OUTPUT(mode="XYZ")
    with pytest.raises(ValueError):
        P.OUTPUT(entities=["SPEED"])
    with pytest.raises(ValueError):
        P.OUTPUT(format="PNG")
    with pytest.raises(ValueError):
        P.OUTPUT(subset=[7])
    with pytest.raises(NotImplementedError):
        P.OUTPUT(format="XDMF")
    with pytest.raises(TypeError):
        P.OUTPUT(entities="NUMBER")


def test_torsion_spring_and_joint_files(tmp_path):
    """ST and JT modes (output.cpp:2969-3197, 4590-4730): a torsion spring
    between two cubes and a ball joint holding one of them, through DEM on
    the oracle; the files carry each item's point, number and vectors as
    the state gives them."""
    from oracle import orc
    from tests import f3_decks as D
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    a = D.cube(mat, 0, 0, 0)
    b = D.cube(mat, 0, 0, 1)
    P.TORSION_SPRING(a, b, (0, 0, 1), (1, 0, 0), kroll=D.K, kpitch=D.K, kyaw=D.K, refpnt=(0.5, 0.5, 1.0))
    P.BALL_JOINT(b, (0.5, 0.5, 2.0))
    P.VELOCITY(a, angular=(0.5, -0.2, 0.1))
    P.GRAVITY(0.0, 0.0, -9.81)
    P.OUTPUT(["NUMBER", "ZDIR", "XDIR", "YDIR", "TRQROT", "TRQTOT", "TRQSPR", "JREAC"], [a, b], mode=["ST", "JT"],
             format=["VTK"])
    P._use_engine(orc.Oracle(1))
    prefix = os.path.join(str(tmp_path), "run")
    # output points after the steps starting at 0.005 and 0.010; the last is
    # the final state
    P.DEM(0.0105, 1e-3, 0.005, prefix=prefix)
    m = P.state()
    st = sorted(glob.glob(prefix + "1st.vtk.*"))
    jt = sorted(glob.glob(prefix + "1jt.vtk.*"))
    assert st and jt
    last = max(st, key=lambda p: int(p.rsplit(".", 1)[1]))
    text = open(last).read()
    # the spring is listed once for each of its particles, as the
    # reference's per-particle maps list it
    assert "PARMEC torsional springs output" in text and "POINTS 2 float" in text
    z = m.trqzdir1[:, 0]
    assert "VECTORS ZDIR float\n%g %g %g\n%g %g %g\n" % (tuple(z) * 2) in text
    L = m.rotation[:, a]
    d = np.array([0.5, 0.5, 1.0]) - m.position[3:, a]
    p = [m.position[c, a] + (L[c] * d[0] + L[3 + c] * d[1] + L[6 + c] * d[2]) for c in range(3)]
    assert text.split("POINTS 2 float\n")[1].split("\n")[0] == "%g %g %g" % tuple(p)
    last = max(jt, key=lambda p: int(p.rsplit(".", 1)[1]))
    text = open(last).read()
    assert "VECTORS JREAC float\n%g %g %g\n" % tuple(m.jreac[:, 0]) in text
    assert "SCALARS number int\nLOOKUP_TABLE default\n0\n" in text
