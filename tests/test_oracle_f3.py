"""Known-answer tests of the oracle's torsion springs, UNSPRING, load
curves, RESTRAIN and BALL_JOINT (CPU): closed forms of the motions they
produce, the reference's own quirks where they decide a value, and the
joint solver against a textbook dense solve.  The GPU path is compared
with the oracle bit for bit in test_gpu_f3.py."""
import math

import numpy as np
import pytest

from oracle import orc
from parmec_amd import parmec as P
from parmec_amd import pmx
from tests import f3_decks as D


def _oracle(m):
    o = orc.Oracle(0)
    o.upload(m.copy())
    return o


def _down(o, m):
    mm = m.copy()
    o.download(mm)
    return mm


# ------------------------------------------------------------ joint solver

def test_skyline_lu_matches_dense_solve():
    """amgcl's skyline LU (profile of non-zeros, Crout, diagonal inverses)
    solves what numpy's LU solves, also with zero runs at the start of rows
    and columns (short skylines) and exact zeros inside the profile."""
    rng = np.random.default_rng(7)
    for n in (3, 6, 9, 15, 30):
        A = rng.normal(size=(n, n)) + n * np.eye(n)
        # short skylines: zero the far corners
        for i in range(n):
            for j in range(n):
                if abs(i - j) > max(1, n // 3) and rng.uniform() < 0.7:
                    A[i, j] = 0.0
        A[n // 2, max(0, n // 2 - 1)] = 0.0   # a zero inside the profile
        b = rng.normal(size=n)
        x = orc.skyline_lu(A, b)
        ref = np.linalg.solve(A, b)
        assert np.abs(x - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_skyline_lu_zero_pivot():
    A = np.array([[0.0, 1.0], [1.0, 0.0]])
    assert orc.skyline_lu(A, np.ones(2)) is None          # D[0] = 0: the reference's precondition
    A = np.array([[1.0, 2.0], [2.0, 4.0]])
    assert orc.skyline_lu(A, np.ones(2)) is None          # singular: a zero pivot later


def test_cuthill_mckee_orders():
    """amgcl's Cuthill-McKee walk (it expands nodes while `node > 0`, so the
    first node is never expanded and the walk restarts at node 1): the dense
    pattern of a joint set keeps the identity order; a sparse pattern whose
    node 1 reaches 3 before 2 orders [0, 1, 3, 2] (derived by hand)."""
    for n in (3, 6, 12):
        ptr = np.arange(n + 1) * n
        col = np.tile(np.arange(n), n)
        assert list(orc.cuthill_mckee(ptr, col)) == list(range(n))
    rows = [[0], [1, 3], [2, 3], [1, 2, 3]]
    ptr = np.cumsum([0] + [len(r) for r in rows])
    assert list(orc.cuthill_mckee(ptr, np.concatenate(rows))) == [0, 1, 3, 2]


# --------------------------------------------------------- torsion springs

def _cube_on_torsion(curves, w, zdir=(1, 0, 0), xdir=(0, 0, 1), **kw):
    P.RESET()
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    p = D.cube(mat, 0, 0, 0)
    P.TORSION_SPRING(p, -1, zdir, xdir, **curves, **kw)
    P.VELOCITY(p, angular=w)
    m = P._build_model()
    P.RESET()
    return m


def test_torsion_roll_is_a_harmonic_oscillator():
    """Roll about xdir (global z here) with a linear restoring curve k: the
    roll angle follows (w0/W) sin(W t), W = sqrt(k/I), I = m/6 for the unit
    cube; pitch and yaw stay zero."""
    k, w0, h = 1E4, 1.0, 1e-4
    m = _cube_on_torsion({"kroll": [-1, k, 1, -k]}, (0, 0, w0))
    I = m.mass[0] / 6.0
    W = math.sqrt(k / I)
    o = _oracle(m)
    t = 0.0
    for _ in range(4):
        o.dem(0.05, h, pmx.forcing())
        t += 0.05
        mm = _down(o, m)
        roll = mm.trqrpy[0, 0]
        assert abs(roll - (w0 / W) * math.sin(W * t)) <= 2e-3 * (w0 / W)
        assert abs(mm.trqrpy[1, 0]) < 1e-12 and abs(mm.trqrpy[2, 0]) < 1e-12
        # the spring torque: -k roll, the axis the current x direction
        assert abs(mm.trqrpyspr[0, 0] + k * roll) <= 1e-9 * k * abs(roll)
        np.testing.assert_allclose(mm.trqxdir1[:, 0], mm.rotation[:, 0].reshape(3, 3, order="F") @ [0, 0, 1],
                                   atol=1e-15)


def test_torsion_critical_ratio_damps_critically():
    """A critical damping ratio 1 gives -2 sqrt(k I) roll-rate damping: the
    roll angle follows w0 t exp(-W t) and never overshoots."""
    k, w0, h = 1E4, 1.0, 1e-4
    m = _cube_on_torsion({"kroll": [-1, k, 1, -k], "droll": 1.0}, (0, 0, w0))
    I = m.mass[0] / 6.0
    W = math.sqrt(k / I)
    o = _oracle(m)
    t = 0.0
    for _ in range(8):
        o.dem(0.1, h, pmx.forcing())
        t += 0.1
        roll = _down(o, m).trqrpy[0, 0]
        exact = w0 * t * math.exp(-W * t)
        assert roll > 0.0
        assert abs(roll - exact) <= 5e-3 * w0 / W


def test_torsion_cone_splits_the_torque_along_the_angles():
    """cone ('roll', 'pitch'): the roll curve acts on sqrt(roll^2 +
    pitch^2) and its torque is split by roll/angle and pitch/angle; the
    stiffness of the pitch is the roll's."""
    k = 2E4
    m = _cube_on_torsion({"kroll": [-1, k, 1, -k], "kyaw": [-1, k, 1, -k]}, (0, 0, 0), cone=("roll", "pitch"))
    m.trqrpy[:, 0] = [0.03, 0.04, 0.0]
    o = _oracle(m)
    o.dem(1e-9, 1e-9, pmx.forcing())
    mm = _down(o, m)
    roll, pitch = mm.trqrpy[0, 0], mm.trqrpy[1, 0]
    angle = math.hypot(roll, pitch)
    assert abs(mm.trqrpyspr[0, 0] - (-k * angle) * roll / angle) <= 1e-12 * k
    assert abs(mm.trqrpyspr[1, 0] - (-k * angle) * pitch / angle) <= 1e-12 * k
    assert mm.trqrpyspr[2, 0] == 0.0


def test_torsion_stale_part2_velocity_quirk():
    """trqspr_task reloads part2's angular velocity only where part2 changes
    and zeroes it for a spring to the ground (forces.ispc:1500-1526): in the
    sorted order (0, 2), (1, -1), (1, 2) the third spring reads zero for
    particle 2's spin, so its roll rate is that of particle 1 alone."""
    P.RESET()
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    a = D.cube(mat, 0, 0, 0)
    b = D.cube(mat, 2, 0, 0)
    c = D.cube(mat, 4, 0, 0)
    for p1, p2 in ((a, c), (b, -1), (b, c)):
        P.TORSION_SPRING(p1, p2, (1, 0, 0), (0, 0, 1), kroll=[-1, 1.0, 1, -1.0])
    P.VELOCITY(b, angular=(0, 0, 0.5))
    P.VELOCITY(c, angular=(0, 0, 2.0))
    m = P._build_model()
    P.RESET()
    h = 1e-6
    o = _oracle(m)
    # the first call's half step leaves the spatial spin scaled by h/2
    # (euler_task, parmec.ispc:260-264); the second call's step reads it whole
    o.dem(h, h, pmx.forcing())
    r0 = _down(o, m).trqrpy[0].copy()
    o.dem(h, h, pmx.forcing())
    mm = _down(o, m)
    d = mm.trqrpy[0] - r0
    # spring order after sorting: (a, c), (b, -1), (b, c)
    assert list(mm.trqsprpart[0]) == [a, b, b] and list(mm.trqsprpart[1]) == [c, -1, c]
    assert abs(d[0] - (0.0 - 2.0) * h) <= 1e-9 * h        # a - c
    assert abs(d[1] - 0.5 * h) <= 1e-9 * h                # b - ground
    assert abs(d[2] - 0.5 * h) <= 1e-9 * h                # b - (zeros, not c)


def test_torsion_zdir_quirk():
    """TORSION_SPRING scales zdir by 1/|zdir|^2 and keeps xdir as given
    (input.cpp:1878-1888)."""
    P.RESET()
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    p = D.cube(mat, 0, 0, 0)
    P.TORSION_SPRING(p, -1, (2, 0, 0), (0, 0, 3))
    m = P._build_model()
    P.RESET()
    np.testing.assert_array_equal(m.trqzdir0[:, 0], [0.5, 0.0, 0.0])
    np.testing.assert_array_equal(m.trqxdir0[:, 0], [0.0, 0.0, 3.0])


# ---------------------------------------------------------------- UNSPRING

def _pulled_spring(nfreq=1, nsteps=1, unload=None, limit=2E3):
    """A cube held still (restrained) and a heavy cube moving away at 1 m/s,
    one linear spring k = 1e6 between them; UNSPRING on its elastic force
    above `limit`."""
    P.RESET()
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    heavy = P.MATERIAL(1E12, 1E9, 0.25)
    a = D.cube(mat, 0, 0, 0)
    b = D.cube(heavy, 1.5, 0, 0)
    P.RESTRAIN(a, [1, 0, 0, 0, 1, 0, 0, 0, 1], [1, 0, 0, 0, 1, 0, 0, 0, 1])
    P.VELOCITY(b, (1.0, 0.0, 0.0))
    s = P.SPRING(a, (1, 0.5, 0.5), b, (1.5, 0.5, 0.5), [-1, -1E6, 1, 1E6])
    kw = {}
    if unload is not None:
        kw["unload"] = P.TSERIES(unload)
    P.UNSPRING([s], [s], (None, limit), nfreq=nfreq, nsteps=nsteps, **kw)
    m = P._build_model()
    P.RESET()
    return m


def _steps(m, n, h=1e-4):
    """n dem calls of one step each: per call the elastic force, the spring
    state and the failures left."""
    o = _oracle(m)
    out = []
    for _ in range(n):
        o.dem(h, h, pmx.forcing())
        mm = _down(o, m)
        out.append((mm.sprfrc[1, 0], int(mm.unspring[0]), int(mm.nfreq[0]), mm.stroke[0, 0]))
    return out


def test_unspring_counts_failures_then_switches_off():
    """The force exceeds the limit from some step on; nfreq = 3 failures
    later the spring drops to zero force for ever (unspring -1)."""
    seq = _steps(_pulled_spring(nfreq=3), 40)
    first = next(i for i, (f, _, _, _) in enumerate(seq) if f > 2E3)
    for i, (f, u, n, _) in enumerate(seq):
        if i < first:
            assert (u, n) == (-2, 3)
        elif i < first + 2:
            assert (u, n) == (-2, 3 - (i - first + 1))
        else:
            assert n == 0 and u == -1
            if i > first + 2:
                assert f == 0.0


def test_unspring_tests_every_nsteps_dem_call():
    """stepnum counts dem() calls (parmec.cpp:2523): with nsteps = 2 the
    condition is tested in the calls whose number is even only."""
    seq = _steps(_pulled_spring(nsteps=2), 40)
    first = next(i for i, (f, _, _, _) in enumerate(seq) if f > 2E3)
    trig = next(i for i, (_, _, n, _) in enumerate(seq) if n == 0)
    assert trig == (first if first % 2 == 0 else first + 1)


def test_unspring_unloading_curve():
    """Switched onto an unloading curve of slope s, the elastic force falls
    by |s ds| per step (forces.ispc:1080-1115) until it would change sign,
    then the spring holds zero for ever."""
    s = 2E7
    seq = _steps(_pulled_spring(unload=[0.0, 0.0, 1.0, s]), 60)
    trig = next(i for i, (_, _, n, _) in enumerate(seq) if n == 0)
    assert seq[trig][1] == 0                             # the curve's load-curve number
    f0 = seq[trig][0]
    for i in range(trig + 1, len(seq)):
        f, u, _, st = seq[i]
        if u == -1:
            assert f == 0.0
            break
        ds = st - seq[i - 1][3]
        exp = seq[i - 1][0] - abs(s * ds)
        assert f == pytest.approx(exp, rel=1e-12, abs=1e-9)
        assert f < f0
    else:
        raise AssertionError("the spring never reached zero force")


# --------------------------------------------------------- stroke offsets

def test_stroke_offset_follows_its_load_curve():
    """A spring with offset = TSERIES: force = curve(stroke - offset(t)) at
    the step's start time (forces.ispc:915-931)."""
    m = D.offset_deck()
    o = _oracle(m)
    h = 1e-4
    t = 0.0
    ramp = lambda t: 0.1 * t
    for _ in range(5):
        o.dem(10 * h, h, pmx.forcing())
        t += 10 * h
        mm = _down(o, m)
        # the last step started at t - h; spring 0 (sorted first) is the ramp's
        st = mm.stroke[0, 0]
        f_el = mm.sprfrc[1, 0]
        assert f_el == pytest.approx(1E6 * (st - ramp(t - h)), rel=1e-9, abs=1e-6)


# -------------------------------------------------------------- restraints

def test_restrained_velocities_closed_form():
    """restrain.py's cubes under a constant gravity: restrained linear
    directions keep zero velocity (the velocity projected once, the force
    every step), the others follow v0 + g t."""
    m = D.restrain_deck()
    o = _oracle(m)
    g = np.array([0.3, -0.2, -10.0])
    T, h = 0.5, 1e-2
    o.dem(T, h, pmx.forcing(gravity=tuple(g)))
    mm = _down(o, m)
    free = np.array([1.0, 1.0, 1.0]) + g * T
    masks = {0: [1, 1, 1], 1: [0, 1, 1], 2: [0, 0, 1], 3: [0, 0, 0]}
    for p, mask in masks.items():
        np.testing.assert_allclose(mm.linear[:, p], free * mask, atol=1e-12)
    for p in (4, 5, 6):
        np.testing.assert_allclose(mm.linear[:, p], free, atol=1e-12)
    # angular restraints: the torque components along them are removed
    for p, dirs in ((4, [0]), (5, [0, 1]), (6, [0, 1, 2])):
        for d in dirs:
            assert mm.torque[d, p] == 0.0
    # the velocity projection at the start of the call
    P.RESET()
    o2 = _oracle(m)
    o2.dem(1e-12, 1e-12, pmx.forcing())
    m2 = _down(o2, m)
    assert m2.angular[3, 4] == pytest.approx(0.0, abs=1e-9)
    np.testing.assert_allclose(m2.angular[3:5, 5], 0.0, atol=1e-9)
    np.testing.assert_allclose(m2.angular[3:, 6], 0.0, atol=1e-9)


# -------------------------------------------------------------- ball joints

def _hanging_cube(offset=(0.0, 0.0), w=(0.0, 0.0, 0.0)):
    P.RESET()
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    p = D.cube(mat, 0, 0, 0)
    P.BALL_JOINT(p, (0.5 + offset[0], 0.5 + offset[1], 1.5))
    P.VELOCITY(p, angular=w)
    P.GRAVITY(0., 0., -10.)
    m = P._build_model()
    P.RESET()
    return m


def test_ball_joint_holds_a_hanging_body():
    """A cube at rest below its joint: the reaction is m g straight up (the
    joint block reduces to 1/m on the vertical), and the body stays put."""
    m = _hanging_cube()
    o = _oracle(m)
    h = 1e-3
    o.dem(h, h, pmx.forcing(gravity=(0.0, 0.0, -10.0)))
    mm = _down(o, m)
    np.testing.assert_allclose(mm.jreac[:, 0], [0.0, 0.0, 10.0 * m.mass[0]], rtol=1e-12, atol=1e-9)
    o.dem(0.5, h, pmx.forcing(gravity=(0.0, 0.0, -10.0)))
    mm = _down(o, m)
    np.testing.assert_allclose(mm.position[:3, 0], [0.5, 0.5, 0.5], atol=1e-9)


def test_ball_joint_pendulum_period():
    """The cube on a joint 1 m above its centre swings with the compound
    pendulum's small-oscillation period 2 pi sqrt(I_p / (m g L)),
    I_p = m/6 + m L^2, and its joint point stays on the fixed point."""
    m = _hanging_cube(w=(0.05, 0.0, 0.0))
    o = _oracle(m)
    g, L, h = 10.0, 1.0, 1e-3
    Ip = m.mass[0] / 6.0 + m.mass[0] * L * L
    T = 2 * math.pi * math.sqrt(Ip / (m.mass[0] * g * L))
    ys, ts = [], []
    for k in range(int(1.2 * T / 0.01)):
        o.dem(0.01, h, pmx.forcing(gravity=(0.0, 0.0, -g)))
        mm = _down(o, m)
        ys.append(mm.position[1, 0] - 0.5)
        ts.append(0.01 * (k + 1))
        # the joint point of the body (reference (0.5, 0.5, 1.5)) stays put,
        # but for the first call's unconstrained half step (w h/2 L)
        L9 = mm.rotation[:, 0].reshape(3, 3, order="F")
        jp = mm.position[:3, 0] + L9 @ (np.array([0.5, 0.5, 1.5]) - mm.position[3:, 0])
        assert np.abs(jp - [0.5, 0.5, 1.5]).max() < 0.05 * h / 2 * L + 1e-6
    ys = np.array(ys)
    # the first two downward zero crossings of y are a period apart
    cross = [ts[i] for i in range(1, len(ys)) if ys[i - 1] > 0 >= ys[i]]
    assert len(cross) >= 1
    s = [ts[i] for i in range(1, len(ys)) if ys[i - 1] < 0 <= ys[i]]
    assert abs((cross[0] - 0.0) - T / 2) < 0.02 * T or (s and abs(s[0] - cross[0] - T / 2) < 0.02 * T)


def test_ball_joint_chain_keeps_its_joints():
    """ball_joint_chain.py: after half a second of swinging, every joint's
    two body points still coincide (the velocity-level constraint drifts
    little), and the sets are the chains of the forest."""
    m = D.chain_deck(6, forest=2, edge=8)
    o = _oracle(m)
    o.dem(0.5, 1e-3, pmx.forcing(gravity=(0.0, 0.0, -10.0)))
    mm = _down(o, m)
    for j in range(mm.jnum):
        pts = []
        for k in range(2):
            p = mm.jpart[k, j]
            if p < 0:
                pts.append(mm.jpoint[:, j])
                continue
            L9 = mm.rotation[:, p].reshape(3, 3, order="F")
            pts.append(mm.position[:3, p] + L9 @ (mm.jpoint[:, j] - mm.position[3:, p]))
        assert np.abs(pts[0] - pts[1]).max() < 1e-4, j
    assert np.abs(mm.jreac).max() > 1e3


# ------------------------------------------------------------ deck checks

def test_deck_rejects_bad_f3_input():
    P.RESET()
    mat = P.MATERIAL(1E3, 1E9, 0.25)
    p = D.cube(mat, 0, 0, 0)
    with pytest.raises(ValueError):
        P.TORSION_SPRING(p, -1, (1, 0, 0), (1, 0, 0))                  # not orthogonal
    with pytest.raises(ValueError):
        P.TORSION_SPRING(p, -1, (1, 0, 0), (0, 0, 1), kroll=[0, 1, 0, 2])   # angles must increase
    with pytest.raises(ValueError):
        P.TORSION_SPRING(p, -1, (1, 0, 0), (0, 0, 1), cone=("roll", "roll"))
    with pytest.raises(ValueError):
        P.TORSION_SPRING(p, -1, (1, 0, 0), (0, 0, 1), cone=("roll", "pitch"))   # no roll curve
    with pytest.raises(ValueError):
        P.RESTRAIN(p, [1, 0])
    with pytest.raises(ValueError):
        P.BALL_JOINT(p, (0, 0, 0), 5)
    s = P.SPRING(p, (0, 0, 0), -1, (0, 0, 1), [-1, -1, 1, 1])
    with pytest.raises(ValueError):
        P.UNSPRING([s], [s], (1.0, 0.0))                                 # lower >= upper
    P.UNSPRING([s], [s], (None, 1.0))
    with pytest.raises(ValueError):
        P.UNSPRING([s], [s], (None, 1.0))                                # already in use
    ts = P.TSERIES(1.0)
    with pytest.raises(ValueError):
        P.SPRING(p, (0, 0, 0), -1, (0, 0, 1), [-1, -1, 1, 1], offset=ts)   # a constant is no load curve
    P.RESET()
