"""Closed-form known answers for every SPRING mode of the oracle
(forces.ispc:641-1412, eqm parmec.ispc:423-482).

Each case takes one step from a state written down here, so every expected
value is a short formula of that state, evaluated with numpy:
  - directions (forces.ispc:857-913): follower z = (q - p)/|q - p|, len =
    |q - p|; constant z = d, len = (q - p).d; planar z = the part of q - p
    normal to n, normalised, len its length; point-to-plane z = n (turned
    with the second body, L_k n), len = (p - q).z;
  - stroke = len - stroke0, F = the curve at the stroke, force on the first
    body z (F + dashpot), on the second minus that, torques a x f and
    -(b x f) about the bodies' centres (1069, 1203-1322);
  - dashpots: a table looked up at the normal velocity z.(u_k - u_j) (1054-
    1063), or the critical ratio c = ratio 2 sqrt(|k| m_eq) with m_eq =
    1/(z^T M z), M = sum over bodies of (1/m) I + A^T J^-1 A (A the skew
    matrix of the arm the springs task passes: the reference position minus
    the body-frame offset of the spring point, i.e. 2 X0 - Y), or the
    extremum eigenvalue of M when z = 0 (the largest when M is diagonal,
    the smallest otherwise, macros.h:1539-1586);
  - the dashpot counts only where the spring force is non-zero (1069);
  - friction (1125-1200): without kskn the tangential velocity's direction
    times mu |total|; with kskn the tangential displacement history
    ds = s0 + u_t h, force kskn (|k| ds + |c| u_t), capped at mu |total| with
    the history rescaled to the cap.
A sign flip, a swapped operand or a dropped term in any of these changes a
checked value by far more than the tolerance.
"""
import math

import numpy as np
import pytest

from parmec_amd import parmec as P
from parmec_amd import pmx

H = 1e-4
K = 2.0e4                      # slope of the linear test curve
LIN = [-1.0, -K, 1.0, K]       # F = K s, extrapolated beyond |s| > 1
RTOL = 1e-11


@pytest.fixture(autouse=True)
def fresh():
    P.RESET()
    yield
    P.RESET()


def bodies(xs, vs=None, r=0.05, rho=1000.0):
    mat = P.MATERIAL(rho, 1e6, 0.2)
    for i, x in enumerate(xs):
        P.SPHERE(tuple(float(c) for c in x), r, mat, 1)
        if vs is not None:
            P.VELOCITY(i, tuple(float(c) for c in vs[i]), (0.0, 0.0, 0.0))


def one_step(m, h=H):
    """The model after one oracle step from rest time 0 (the first call
    first moves x by h/2 v, euler, parmec.ispc:224-272)."""
    from oracle import orc
    o = orc.Oracle(1)
    mm = m.copy()
    o.upload(mm)
    o.dem(0.5 * h, h, pmx.forcing())
    o.download(mm)
    o.close()
    return mm


def state(m, h=H):
    """Positions of the step (after the half-step move) and velocities."""
    x = m.position[:3].T + 0.5 * h * m.linear.T
    return x, m.linear.T.copy()


def displace(m, i, d):
    m.position[:3, i] += d
    m.center[:3, i] += d


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = max(float(np.abs(b).max()), 1e-300)
    assert float(np.abs(a - b).max()) <= rtol * scale, (a, b)


def unit(v):
    return v / np.linalg.norm(v)


def check_pair(mm, z, total, fric, p, q, x):
    f = z * total + fric
    close(mm.force[:, 0], f)
    close(mm.force[:, 1], -f)
    close(mm.torque[:, 0], np.cross(p - x[0], f))
    close(mm.torque[:, 1], -np.cross(q - x[1], f))


def test_follower_direction():
    bodies([(0.0, 0.0, 0.0), (0.3, 0.4, 0.0)])
    P.SPRING(0, (0.0, 0.0, 0.0), 1, (0.3, 0.4, 0.0), LIN)
    m = P._build_model()
    displace(m, 1, np.array([0.03, 0.04, -0.02]))
    mm = one_step(m)
    x, _ = state(m)
    w = x[1] - x[0]
    z = unit(w)
    s = np.linalg.norm(w) - 0.5
    close(mm.stroke[0, 0], s)
    close(mm.sprdir[:3, 0], z)
    close(mm.sprfrc[:2, 0], [K * s, K * s])
    check_pair(mm, z, K * s, 0.0, x[0], x[1], x)


def test_constant_direction():
    d = np.array([1.0, 2.0, 2.0]) / 3.0
    bodies([(0.0, 0.0, 0.0), (0.3, 0.1, 0.2)])
    P.SPRING(0, (0.0, 0.0, 0.0), 1, (0.3, 0.1, 0.2), LIN, direction=(1.0, 2.0, 2.0))
    m = P._build_model()
    displace(m, 1, np.array([0.02, -0.01, 0.03]))
    mm = one_step(m)
    x, _ = state(m)
    s0 = np.array([0.3, 0.1, 0.2]) @ d
    s = (x[1] - x[0]) @ d - s0
    close(mm.stroke[0, 0], s)
    check_pair(mm, d, K * s, 0.0, x[0], x[1], x)


def test_planar_direction():
    n = np.array([0.0, 0.6, 0.8])
    bodies([(0.0, 0.0, 0.0), (0.4, 0.2, -0.1)])
    P.SPRING(0, (0.0, 0.0, 0.0), 1, (0.4, 0.2, -0.1), LIN, direction=tuple(n), planar="ON")
    m = P._build_model()
    displace(m, 1, np.array([0.01, 0.05, 0.02]))
    mm = one_step(m)
    x, _ = state(m)

    def inplane(w):
        return w - (w @ n) * n
    w = inplane(x[1] - x[0])
    z = unit(w)
    s = np.linalg.norm(w) - np.linalg.norm(inplane(np.array([0.4, 0.2, -0.1])))
    close(mm.stroke[0, 0], s)
    close(mm.sprdir[:3, 0], z)
    check_pair(mm, z, K * s, 0.0, x[0], x[1], x)


def test_point_to_fixed_plane():
    """Body 0 against the plane z = -1 (normal +z): len = (p - q).n, q the
    projection of the spring point; only the first body is loaded."""
    n = np.array([0.0, 0.0, 1.0])
    bodies([(0.1, 0.2, 0.3)])
    P.SPRING(0, (0.1, 0.2, 0.3), -1, [(5.0, -2.0, -1.0), tuple(n)], LIN)
    m = P._build_model()
    displace(m, 0, np.array([0.05, -0.02, 0.05]))
    mm = one_step(m)
    x, _ = state(m)
    q = np.array([0.1, 0.2, -1.0])
    s = (x[0] - q) @ n
    close(mm.stroke[0, 0], s)
    f = n * K * s
    close(mm.force[:, 0], f)
    close(mm.torque[:, 0], np.zeros(3))           # the spring point is the centre


def test_point_to_plane_turning_with_second_body():
    """Point-to-plane with a second body: the plane normal and the plane
    point are turned by the second body's rotation (forces.ispc:813-836)."""
    n = unit(np.array([0.2, -0.3, 1.0]))
    bodies([(0.0, 0.0, 0.5), (0.1, 0.2, -0.3)])
    P.SPRING(0, (0.0, 0.0, 0.5), 1, [(0.1, 0.2, -0.3), tuple(n)], LIN)
    m = P._build_model()
    c, s_ = math.cos(0.3), math.sin(0.3)
    R = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
    m.rotation[:, 1] = R.T.reshape(-1)            # column-major: element (r, c) at 3c + r
    mm = one_step(m)
    x, _ = state(m)
    p1 = np.array([0.0, 0.0, 0.5])
    X1 = np.array([0.1, 0.2, -0.3])
    Y1 = p1 - ((p1 - X1) @ n) * n                 # the plane point: p1 projected
    q = x[1] + R @ (Y1 - X1)
    z = R @ n
    s = (x[0] - q) @ z
    close(mm.stroke[0, 0], s)
    close(mm.sprdir[:3, 0], z)
    check_pair(mm, z, K * s, 0.0, x[0], q, x)


@pytest.mark.parametrize("curve,expect_dashpot", [(LIN, True), ([-1.0, 0.0, 1.0, 0.0], False)])
def test_table_dashpot(curve, expect_dashpot):
    """Dashpot table F_d = 5 v; with a zero spring force the total is zero
    whatever the velocity (the reference adds the dashpot only where the
    spring force is non-zero, forces.ispc:1069)."""
    bodies([(0.0, 0.0, 0.0), (0.5, 0.0, 0.0)], vs=[(-0.1, 0.02, 0.0), (0.2, 0.0, 0.01)])
    P.SPRING(0, (0.0, 0.0, 0.0), 1, (0.5, 0.0, 0.0), curve, dashpot=[-10.0, -50.0, 10.0, 50.0])
    m = P._build_model()
    displace(m, 1, np.array([0.01, 0.0, 0.0]))
    mm = one_step(m)
    x, v = state(m)
    w = x[1] - x[0]
    z = unit(w)
    s = np.linalg.norm(w) - 0.5
    fs = K * s if expect_dashpot else 0.0
    vel = z @ (v[1] - v[0])
    total = fs + (5.0 * vel if expect_dashpot else 0.0)
    close(mm.sprfrc[:2, 0], [total, fs])
    check_pair(mm, z, total, 0.0, x[0], x[1], x)


def m_eq_matrix(m, arms):
    M = np.zeros((3, 3))
    for i, A in arms:
        j = m.inertia[0, i]
        M += np.eye(3) / m.mass[i] + ((A @ A) * np.eye(3) - np.outer(A, A)) / j
    return M


def test_critical_dashpot_two_bodies():
    """Critical damping with the equivalent mass of two bodies whose spring
    points are off their centres; the arms are those the springs task
    passes to eqm (2 X0 - Y, see the module docstring)."""
    X = [np.array([0.0, 0.0, 0.0]), np.array([0.5, 0.0, 0.3])]
    Y = [np.array([0.0, 0.02, 0.0]), np.array([0.5, 0.0, 0.0])]
    bodies(X, vs=[(0.0, 0.0, 0.0), (0.2, 0.0, 0.0)])
    zeta = 0.3
    P.SPRING(0, tuple(Y[0]), 1, tuple(Y[1]), LIN, dashpot=zeta)
    m = P._build_model()
    displace(m, 1, np.array([0.01, 0.0, 0.0]))
    mm = one_step(m)
    x, v = state(m)
    p = x[0] + (Y[0] - X[0])
    q = x[1] + (Y[1] - X[1])
    w = q - p
    z = unit(w)
    s = np.linalg.norm(w) - np.linalg.norm(Y[1] - Y[0])
    M = m_eq_matrix(m, [(0, 2 * X[0] - Y[0]), (1, 2 * X[1] - Y[1])])
    meq = 1.0 / (z @ M @ z)
    c = zeta * 2.0 * math.sqrt(K * meq)
    total = K * s + c * (z @ (v[1] - v[0]))
    close(mm.sprfrc[:2, 0], [total, K * s])
    check_pair(mm, z, total, 0.0, p, q, x)


def test_friction_without_stick():
    """kskn = 0: friction mu |total| along the tangential relative velocity."""
    mu = 0.4
    bodies([(0.0, 0.0, 0.0), (0.5, 0.0, 0.0)], vs=[(0.0, 0.0, 0.0), (0.05, 0.3, 0.1)])
    P.SPRING(0, (0.0, 0.0, 0.0), 1, (0.5, 0.0, 0.0), LIN, friction=mu)
    m = P._build_model()
    displace(m, 1, np.array([0.02, 0.0, 0.0]))
    mm = one_step(m)
    x, v = state(m)
    z = unit(x[1] - x[0])
    s = np.linalg.norm(x[1] - x[0]) - 0.5
    u = v[1] - v[0]
    ut = u - (z @ u) * z
    fr = mu * abs(K * s) * unit(ut)
    close(mm.sprfrc[2, 0], np.linalg.norm(fr))
    check_pair(mm, z, K * s, fr, x[0], x[1], x)


@pytest.mark.parametrize("speed,sticks", [(0.01, True), (10.0, False)])
def test_friction_stick_slip(speed, sticks):
    """kskn > 0, from an empty history: ds = u_t h, f = kskn K ds (no
    dashpot: its coefficient is 0); below mu |F| it sticks (history ds),
    above it slips (force and history scaled to the cap)."""
    mu, kskn = 0.3, 0.5
    bodies([(0.0, 0.0, 0.0), (0.5, 0.0, 0.0)], vs=[(0.0, 0.0, 0.0), (0.0, speed, 0.5 * speed)])
    P.SPRING(0, (0.0, 0.0, 0.0), 1, (0.5, 0.0, 0.0), LIN, friction=mu, kskn=kskn)
    m = P._build_model()
    displace(m, 1, np.array([0.0005, 0.0, 0.0]))
    mm = one_step(m)
    x, v = state(m)
    z = unit(x[1] - x[0])
    F = K * (np.linalg.norm(x[1] - x[0]) - 0.5)
    u = v[1] - v[0]
    ut = u - (z @ u) * z
    ds = ut * H
    c0 = kskn * K
    fr = c0 * ds
    flim = mu * abs(F)
    assert (np.linalg.norm(fr) <= flim) == sticks
    if sticks:
        hist = ds
    else:
        fr = fr * (flim / np.linalg.norm(fr))
        hist = ds * (flim / (c0 * np.linalg.norm(ds)))
    close(mm.sprsdsp[:, 0], hist)
    check_pair(mm, z, F, fr, x[0], x[1], x)


@pytest.mark.parametrize("offset_y,diagonal", [(0.0, True), (0.1, False)])
def test_zero_length_spring_eigen_mass(offset_y, diagonal):
    """A follower spring of zero length: z = 0, so eqm takes the extremum
    eigenvalue of M (the largest when M is diagonal, the smallest
    otherwise).  Only stick-slip friction reads the resulting damping
    (c1 = kskn |c|, c = ratio 2 sqrt(|k| m_eq)).  Body 1 starts displaced by
    -h/2 v (dyadic numbers: every sum below is exact), so after the first
    half-step move both spring points coincide while the bodies still move
    relative to each other."""
    mu, kskn, zeta = 0.5, 0.4, 0.2
    h = 2.0 ** -13
    v1 = np.array([0.0, 0.125, 0.0])
    Y = np.array([0.25, offset_y, 0.0])
    X = [np.zeros(3), np.array([0.5, 0.0, 0.0])]
    bodies(X, vs=[(0.0, 0.0, 0.0), tuple(v1)])
    P.SPRING(0, tuple(Y), 1, tuple(Y), [-1.0, 10.0, 1.0, 30.0], dashpot=zeta, friction=mu, kskn=kskn)
    m = P._build_model()
    displace(m, 1, -0.5 * h * v1)
    mm = one_step(m, h)
    x, v = state(m, h)
    assert np.array_equal(x[1], X[1])
    M = m_eq_matrix(m, [(0, 2 * X[0] - Y), (1, 2 * X[1] - Y)])
    offdiag = M[1, 0] ** 2 + M[2, 0] ** 2 + M[2, 1] ** 2
    assert (offdiag == 0.0) == diagonal
    lam = np.diag(M).max() if diagonal else np.linalg.eigvalsh(M).min()
    stiff, F = 10.0, 20.0
    c = zeta * 2.0 * math.sqrt(stiff / lam)
    ut = v[1] - v[0]                              # z = 0: all of it is tangential
    fr = kskn * stiff * (ut * h) + kskn * c * ut
    flim = mu * F
    if np.linalg.norm(fr) > flim:
        fr = fr * (flim / np.linalg.norm(fr))
    close(mm.sprfrc[:2, 0], [F, F])
    close(mm.force[:, 0], fr, rtol=1e-9)
    close(mm.force[:, 1], -fr, rtol=1e-9)
