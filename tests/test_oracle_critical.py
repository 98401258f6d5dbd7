"""CPU pins of the oracle's critical and adaptive time steps.

- sin, cos, acos (the header the oracle shares with the device) against
  libm within 2 ulp;
- the 3x3 extremum eigenvalue against numpy, and the reference's diagonal
  branch (the largest entry even when the smallest is asked for);
- time_step_task against the generalized eigenproblem K x = w^2 J x solved
  by scipy (no Cholesky, no trigonometric eigenvalues);
- CRITICAL's per-particle estimate of a sphere resting on a floor and of a
  colliding pair (master side finite, slave side NaN: the slave's kcur is
  never written by the reference, 0 here), and of spring-loaded particles;
  CRITICAL's per-spring estimate, against closed forms;
- the adaptive step: the step length stays the given one while nothing
  loads a particle, equals adaptive x the contact estimate of the previous
  state while in contact (closed form from the sphere's height), and keeps
  its last value after separation; for a spring, the first step's length
  from the spring's stiffness, dashpot slope and arm.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import orc
from parmec_amd import parmec, pmx, workloads


def _ulps(a, b):
    a = np.asarray(a, dtype=np.float64).view(np.int64)
    b = np.asarray(b, dtype=np.float64).view(np.int64)
    a = np.where(a < 0, np.int64(-2**63) - a, a)
    b = np.where(b < 0, np.int64(-2**63) - b, b)
    return np.abs(a - b)


def test_trig_within_two_ulp_of_libm():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-4, 4, 200000), rng.uniform(-1e4, 1e4, 200000), rng.uniform(-0.2, 0.2, 50000),
                        np.arange(-64, 65) * (math.pi / 4)])
    assert _ulps(orc.trig(0, x), np.sin(x)).max() <= 2
    assert _ulps(orc.trig(1, x), np.cos(x)).max() <= 2
    y = np.concatenate([rng.uniform(-1, 1, 200000), 1 - np.ldexp(rng.uniform(0, 1, 2000), -rng.integers(1, 50, 2000)),
                        -1 + np.ldexp(rng.uniform(0, 1, 2000), -rng.integers(1, 50, 2000)),
                        [-1.0, -0.5, 0.0, 0.5, 1.0]])
    assert _ulps(orc.trig(2, y), np.arccos(y)).max() <= 2
    assert np.isnan(orc.trig(2, np.array([1.5, np.nan]))).all()
    assert np.isnan(orc.trig(0, np.array([np.inf]))).all()


def test_extremum_eigenvalue_against_numpy():
    rng = np.random.default_rng(5)
    for _ in range(300):
        B = rng.normal(size=(3, 3))
        A = B + B.T
        w = np.linalg.eigvalsh(A)
        a = A.reshape(9, order="F")
        assert abs(orc.extremum_eigenvalue(a, 0) - w[-1]) < 1e-12 * np.abs(w).max()
        assert abs(orc.extremum_eigenvalue(a, 1) - w[0]) < 1e-12 * np.abs(w).max()
    D = np.diag([3.0, 1.0, 2.0]).reshape(9, order="F")
    assert orc.extremum_eigenvalue(D, 0) == 3.0
    assert orc.extremum_eigenvalue(D, 1) == 3.0   # macros.h:1545-1553: the maximum either way


def _step_closed(nxyz, k, e, m, K, J, mrot):
    wt = math.sqrt(nxyz * k / m)
    lam = scipy.linalg.eigh(K, J, eigvals_only=True).max()
    wr = math.sqrt(lam)
    rt, rr = e / (2 * m * wt), e / (2 * mrot * wr)
    ht = 2 / wt * (math.sqrt(1 + rt * rt) - rt)
    hr = 2 / wr * (math.sqrt(1 + rr * rr) - rr)
    return min(ht, hr), max(wt, wr), max(rt, rr)


def test_time_step_against_generalized_eigenproblem():
    rng = np.random.default_rng(7)
    for _ in range(200):
        B = rng.normal(size=(3, 3))
        J = B @ B.T + 3 * np.eye(3)
        arms = rng.normal(size=(4, 3))
        S = sum(np.dot(a, a) * np.eye(3) - np.outer(a, a) for a in arms)
        krot = [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]]
        kact = rng.uniform(0.1, 2.0, 3)
        k, e, m = rng.uniform(1e3, 1e6), rng.uniform(0.0, 50.0), rng.uniform(0.5, 5.0)
        h, o, r = orc.time_step(kact, k, e, krot, m, J.reshape(9, order="F"))
        he, oe, re = _step_closed(kact.max(), k, e, m, k * S, J, np.linalg.eigvalsh(J)[0])
        assert abs(h - he) < 1e-10 * he and abs(o - oe) < 1e-10 * oe and abs(r - re) < 1e-9 * max(re, 1e-300)
    h, _, _ = orc.time_step([0.0, 0.0, 0.0], 1e6, 1.0, [1, 1, 1, 0, 0, 0], 1.0, np.eye(3).reshape(9))
    assert h == 1.7976931348623157e308   # nothing loads the particle


def _floor(m, size=1.0):
    tris = workloads.box_triangles((0.0, 0.0, 0.0), (size, size, 1.0))[:2]
    m.trinum = 2
    m.tri = np.ascontiguousarray(tris.T)
    m.tricol = np.full(2, 2, dtype=np.int32)
    m.triobs = np.full(2, -1, dtype=np.int32)


def _one_sphere(r, rho, x, v=(0.0, 0.0, 0.0)):
    m = workloads.box_lattice(n=1, dims=(1, 1, 1), radius=r, density=rho, jitter=0.0)
    m.position[:3, 0] = m.position[3:, 0] = x
    m.center[:3, 0] = m.center[3:, 0] = x
    m.linear[:, 0] = v
    return m


def _estimate(wt, wr, e, m, mrot):
    """time_step_task's closed form from the two frequencies; returns the
    estimate and the relative tolerance its evaluation allows (sqrt(1+x^2) -
    x loses about x^2 ulps to cancellation at damping ratio x)"""
    rt, rr = e / (2 * m * wt), e / (2 * mrot * wr)
    h = min(2 / wt * (math.sqrt(1 + rt * rt) - rt), 2 / wr * (math.sqrt(1 + rr * rr) - rr))
    return h, 1e-13 * (1.0 + max(rt, rr) ** 2)


def _floor_estimate(kn, zeta, m, J0, z):
    """one sphere on a floor, contact point straight below at height 0:
    kact = (0, 0, 1), K = kn diag(z^2, z^2, 0), mrot = J0.  The sphere sits
    where the closest point on the floor is exact (x = 1.25, y = 2.5 on a 4 m
    square), so K's off-diagonal terms are exact zeros and the eigenvalue is
    taken from the diagonal; off-diagonal rounding of 1e-32 would send the
    doubly repeated eigenvalue through the trigonometric branch, whose error
    grows like sqrt(eps) there"""
    en = zeta * 2.0 * math.sqrt(kn * m)
    return _estimate(math.sqrt(kn / m), math.sqrt(kn * z * z / J0), en, m, J0)


def test_critical_sphere_on_floor_closed_form():
    r, rho, kn, zeta = 0.5, 1000.0, 1e8, 0.4
    m = _one_sphere(r, rho, (1.25, 2.5, r * 0.999))
    _floor(m, 4.0)
    m.iparam[pmx.SPRING, 0], m.iparam[pmx.DAMPER, 0] = kn, zeta
    o = orc.Oracle(1)
    o.upload(m)
    h0, _, _ = o.critical(pmx.CRIT_PARTICLE)
    assert h0[0] == 1.7976931348623157e308      # before any step: no contact loads it
    o.dem(1e-5, 1e-5, pmx.forcing(gravity=(0.0, 0.0, -9.81)))
    mm = m.copy()
    o.download(mm)
    assert len(o.contacts()) == 1
    h, w, q = o.critical(pmx.CRIT_PARTICLE)
    J0 = m.inertia[0, 0]
    expect, tol = _floor_estimate(kn, zeta, m.mass[0], J0, mm.position[2, 0])
    assert abs(h[0] - expect) < tol * expect
    assert abs(w[0] - math.sqrt(kn * mm.position[2, 0] ** 2 / J0)) < 1e-13 * w[0]


def test_critical_colliding_pair_master_and_slave():
    r, rho, kn, zeta = 0.5, 1000.0, 1e8, 0.2
    m = workloads.box_lattice(n=2, dims=(2, 1, 1), radius=r, density=rho, spacing_factor=1.99, jitter=0.0,
                              spring=kn, damper=zeta, friction=0.0)
    m.trinum = 0
    m.tri = np.zeros((9, 0))
    m.tricol = np.zeros(0, dtype=np.int32)
    m.triobs = np.zeros(0, dtype=np.int32)
    x0 = m.position[:3].copy()
    o = orc.Oracle(1)
    o.upload(m)
    o.dem(1e-6, 1e-6, pmx.forcing())
    mm = m.copy()
    o.download(mm)
    assert len(o.contacts()) == 1
    h, w, q = o.critical(pmx.CRIT_PARTICLE)
    assert np.isnan(h[1])      # slave: kact > 0 but kmax = 0
    mass, J0 = m.mass[0], m.inertia[0, 0]
    en = zeta * 2.0 * math.sqrt(kn * mass / 2.0)
    a0 = 0.5 * (x0[0, 0] + x0[0, 1]) - mm.position[0, 0]
    expect, tol = _estimate(math.sqrt(kn / mass), math.sqrt(kn * a0 * a0 / J0), en, mass, J0)
    assert abs(h[0] - expect) < tol * expect


def _spring_deck(offset=0.2, k=2e6, dash=None, preload=0.0):
    """two spheres joined by a linear spring (force preload + k stroke) from
    an off-centre point of sphere 0 to the centre of sphere 1; dashpot as
    given"""
    parmec.RESET()
    mat = parmec.MATERIAL(1000.0, 1e6, 0.25)
    parmec.SPHERE((0.0, 0.0, 0.0), 0.5, mat, 1)
    parmec.SPHERE((2.0, 0.0, 0.0), 0.5, mat, 1)
    parmec.SPRING(0, (0.0, offset, 0.0), 1, (2.0, 0.0, 0.0), [-10.0, preload - 10 * k, 10.0, preload + 10 * k],
                  dashpot=dash)
    parmec._use_engine(orc.Oracle(1))
    return parmec.state()


def test_critical_perspring_closed_form():
    zeta, k = 0.3, 2e6
    m = _spring_deck(offset=0.0, dash=zeta)
    (hs, j, w, q), = parmec.CRITICAL(perspring=1)
    meq = 1.0 / (1.0 / m.mass[0] + 1.0 / m.mass[1])
    we = math.sqrt(k / meq)
    assert j == 0 and abs(w - we) < 1e-13 * we and q == zeta
    assert abs(hs - 2 / we * (math.sqrt(1 + zeta * zeta) - zeta)) < 1e-13 * hs
    c = 5e4    # a dashpot curve: ratio = slope / (2 meq w)
    _spring_deck(offset=0.0, dash=[-1.0, -c, 1.0, c])
    (hs, j, w, q), = parmec.CRITICAL(perspring=1)
    assert abs(q - c / (2 * meq * we)) < 1e-13 * q


def test_critical_perparticle_springs_closed_form():
    k, c, off = 2e6, 5e4, 0.2
    m = _spring_deck(offset=off, dash=[-1.0, -c, 1.0, c])
    out = parmec.CRITICAL(perparticle=2)
    assert [i for _, i, _, _ in out] == [0, 1] and math.isnan(out[1][0])   # particle 1: arm 0, no rotation term
    d = np.array([2.0, -off, 0.0])
    d /= np.linalg.norm(d)
    mass, J0 = m.mass[0], m.inertia[0, 0]
    expect, tol = _estimate(math.sqrt(np.abs(d).max() * k / mass), math.sqrt(k * off * off / J0), c, mass, J0)
    assert abs(out[0][0] - expect) < tol * expect
    assert parmec.CRITICAL() == out[0][0]
    with pytest.raises(ValueError):
        parmec.CRITICAL(perparticle=3)


def test_adaptive_step_follows_the_contact_estimate():
    r, rho, kn, zeta, ratio, h = 0.5, 1000.0, 1e8, 0.4, 0.2, 2e-4
    m = _one_sphere(r, rho, (1.25, 2.5, r * 1.002), v=(0.0, 0.0, -1.0))
    _floor(m, 4.0)
    m.iparam[pmx.SPRING, 0], m.iparam[pmx.DAMPER, 0] = kn, zeta
    o = orc.Oracle(1)
    o.upload(m)
    steps, z = [], []

    def on_step(k, t, h0, inp):
        inp.gravity[:] = (0.0, 0.0, -9.81)
        inp.output = 1
        steps.append(h0)

    def on_output(k):
        s = m.copy()
        o.download(s)
        z.append(s.position[2, 0])

    o.dem_adaptive(0.05, h, ratio, pmx.forcing(on_step=on_step, on_output=on_output))
    J0, mass = m.inertia[0, 0], m.mass[0]
    first = next(i for i in range(1, len(steps)) if steps[i] != h)
    assert all(s == h for s in steps[:first])
    # the length step i receives was estimated during step i-1, from the
    # state after step i-2
    assert z[first - 3] > r and z[first - 2] < r
    for i in range(first, len(steps)):
        zi = z[i - 2]
        if zi >= r:
            break
        expect, tol = _floor_estimate(kn, zeta, mass, J0, zi)
        assert abs(steps[i] - ratio * expect) < tol * ratio * expect
    sep = next(i for i in range(first, len(z)) if z[i] > r)
    assert all(s == steps[sep + 1] for s in steps[sep + 1:])   # nothing loads it: the last step is kept


def test_adaptive_first_step_from_a_spring():
    k, c, off, ratio, h = 2e6, 5e4, 0.2, 0.5, 1e-3
    # a spring at zero force adds nothing to the adaptive sums (forces.ispc:
    # 1249): the preload makes it active from the first step
    m = _spring_deck(offset=off, dash=[-1.0, -c, 1.0, c], preload=100.0)
    o = orc.Oracle(1)
    o.upload(m)
    steps = []
    o.dem_adaptive(3 * h, h, ratio, pmx.forcing(on_step=lambda k_, t, h0, inp: steps.append(h0)))
    d = np.array([2.0, -off, 0.0])
    d /= np.linalg.norm(d)
    mass, J0 = m.mass[0], m.inertia[0, 0]
    expect, tol = _estimate(math.sqrt(np.abs(d).max() * k / mass), math.sqrt(k * off * off / J0), c, mass, J0)
    assert steps[0] == h and abs(steps[1] - ratio * expect) < tol * ratio * expect
