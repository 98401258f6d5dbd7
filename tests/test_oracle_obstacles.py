"""Known answers for moving obstacles (obstacles.ispc:28-78, the obstacle
branch of the contact law forces.ispc:265-280 and obstaclev
input.cpp:4199-4244), checked on the oracle."""
import math

import numpy as np

from parmec_amd import pmx
from test_oracle_kat import _one_sphere, _oracle_run


def _plate(m, lo, hi, z, color=2):
    """Add a horizontal 2-triangle plate at height z as obstacle 0."""
    tris = np.array([[lo[0], lo[1], z, hi[0], lo[1], z, hi[0], hi[1], z],
                     [lo[0], lo[1], z, hi[0], hi[1], z, lo[0], hi[1], z]])
    m.trinum = 2
    m.tri = np.ascontiguousarray(tris.T)
    m.tricol = np.full(2, color, dtype=np.int32)
    m.triobs = np.full(2, -2, dtype=np.int32)
    m.set_obstacles(1)
    m.trirng[0] = (0, 2)
    m.obspnt[0] = (0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), z)
    return m


def _table(nsteps, ang, lin, obsnum=1):
    row = np.array(list(ang) + list(lin))
    return np.tile(row, (nsteps, obsnum, 1)), np.tile(row, (obsnum, 1))


def test_obstacle_translates_and_rotates_rigidly():
    """No contact: after the first call's obstacles(h) and K steps of
    obstacles(h) the plate has turned by (K+1) h w about its point and moved
    by (K+1) h u (Rodrigues rotation about z)."""
    h, K = 1e-3, 300
    w, u = np.array([0.0, 0.0, 2.0]), np.array([0.1, -0.2, 0.05])
    m = _one_sphere(x=(5.0, 5.0, 5.0))
    m = _plate(m, (-1.0, -0.5, 0), (1.0, 0.5, 0), 0.0)
    p0, t0 = m.obspnt[0].copy(), m.tri.copy()
    tab, tab0 = _table(K, w, u)
    mm, o = _oracle_run(m, (K - 0.5) * h, h, obstab=tab, obstab0=tab0)
    assert o.stats()["steps"] == K
    n = K + 1
    ang = n * h * w[2]
    Rz = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    # point moves by n h u; each step rotates about the point before it moves
    pts = t0.reshape(3, 3, -1)   # [vertex][component][triangle]
    exp = np.empty_like(pts)
    for t in range(pts.shape[2]):
        for v in range(3):
            x = pts[v, :, t] - p0
            # rotations about the moving point compose to Rz about the final point
            exp[v, :, t] = Rz @ x + p0 + n * h * u
    np.testing.assert_allclose(mm.obspnt[0], p0 + n * h * u, atol=1e-13)
    np.testing.assert_allclose(mm.tri.reshape(3, 3, -1), exp, atol=1e-10)
    np.testing.assert_allclose(mm.obsang[0], w, atol=0)
    np.testing.assert_allclose(mm.obslin[0], u, atol=0)


def test_sphere_on_belt_reaches_two_sevenths():
    """A sphere at rest on a belt moving at u: kinetic friction accelerates it
    and spins it backwards until the contact point moves with the belt, at
    v = 2u/7 (translation a = mu g, rotation -5/2 mu g / r)."""
    r, rho, kn, g, mu, u = 0.01, 2000.0, 1e6, 9.81, 0.3, 0.5
    m = _one_sphere(r=r, rho=rho, x=(0.0, -0.5, r - 9.81 * (4.0 / 3.0) * math.pi * r ** 3 * rho / kn))
    m = _plate(m, (-5.0, -1.0, 0), (5.0, 1.0, 0), 0.0)
    m.iparam[pmx.SPRING, 0], m.iparam[pmx.DAMPER, 0] = kn, 0.5
    m.iparam[pmx.FRISTAT, 0] = m.iparam[pmx.FRIDYN, 0] = mu
    h = 0.05 * 2.0 / math.sqrt(kn / m.mass[0])
    t_slip = 2.0 * u / (7.0 * mu * g)
    T = 3.0 * t_slip
    K = int(T / h) + 2
    tab, tab0 = _table(K, (0.0, 0.0, 0.0), (u, 0.0, 0.0))
    mm, o = _oracle_run(m, T, h, gravity=(0.0, 0.0, -g), obstab=tab, obstab0=tab0)
    assert abs(mm.linear[0, 0] - 2.0 / 7.0 * u) < 0.02 * u
    # contact point moves with the belt: v - r omega_y = u
    assert abs(mm.linear[0, 0] - r * mm.angular[4, 0] - u) < 0.02 * u
    assert len(o.contacts()) == 1


def test_rising_plate_carries_sphere():
    """A plate moving up at w carries a resting sphere with it: after the
    transient the sphere moves up at w and stays at the static deflection."""
    r, rho, kn, g, wz = 0.01, 2000.0, 1e6, 9.81, 0.05
    m = _one_sphere(r=r, rho=rho, x=(0.3, -0.2, r - 9.81 * (4.0 / 3.0) * math.pi * r ** 3 * rho / kn))
    m = _plate(m, (-1.0, -1.0, 0), (1.0, 1.0, 0), 0.0)
    m.iparam[pmx.SPRING, 0], m.iparam[pmx.DAMPER, 0] = kn, 0.7
    h = 0.05 * 2.0 / math.sqrt(kn / m.mass[0])
    T = 0.05
    K = int(T / h) + 2
    tab, tab0 = _table(K, (0.0, 0.0, 0.0), (0.0, 0.0, wz))
    mm, o = _oracle_run(m, T, h, gravity=(0.0, 0.0, -g), obstab=tab, obstab0=tab0)
    assert abs(mm.linear[2, 0] - wz) < 1e-3 * wz
    gap = mm.position[2, 0] - mm.obspnt[0, 2]
    assert abs((r - gap) - mm.mass[0] * g / kn) < 1e-3 * mm.mass[0] * g / kn
    assert len(o.contacts()) == 1
