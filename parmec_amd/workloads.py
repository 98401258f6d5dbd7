"""Synthetic Parmec workloads, built directly as SoA models.

Every sphere gets exactly the arrays SPHERE() writes (input.cpp:749-837):
mass = (4/3)*pi*r*r*r*density, inertia = 0.4*m*r*r on the diagonal, identity
rotation, reference position = current position, radii = (r, -1, -1).
Obstacle triangles get triobs = -1 (static OBSTACLE, input.cpp:1367-1437).

box_lattice   : equal spheres on a jittered cubic lattice in a 6-plane box
                (the 10,000-sphere plain FP64 case; any count).
hopper        : polydisperse spheres (lognormal radius, clipped) poured with
                random velocities into a triangulated conical hopper with 60
                degree walls and an open outlet (the 2,000,000-sphere case).
"""
import math

import numpy as np

from . import pmx


def _spheres(model, centers, radii, density, color=1):
    n = len(radii)
    model.position[:3] = centers.T
    model.position[3:] = centers.T
    model.center[:3] = centers.T
    model.center[3:] = centers.T
    model.radii[0] = radii
    model.radii[1] = -1.0
    model.radii[2] = -1.0
    model.part[:] = np.arange(n, dtype=np.int32)
    model.ellcol[:] = color
    mass = (4.0 / 3.0) * math.pi * radii * radii * radii * density
    model.mass[:] = mass
    model.invm[:] = 1.0 / mass
    j = 0.4 * mass * radii * radii
    for k in (0, 4, 8):
        model.inertia[k] = j
        model.inverse[k] = 1.0 / j


def _granular(model, spring, damper, friction, kskn=0.5):
    model.ikind[0] = pmx.GRANULAR_FORCE
    model.iparam[:, 0] = 0.0
    model.iparam[pmx.SPRING, 0] = spring
    model.iparam[pmx.DAMPER, 0] = damper
    model.iparam[pmx.FRISTAT, 0] = friction
    model.iparam[pmx.FRIDYN, 0] = friction
    model.iparam[pmx.KSKN, 0] = kskn


def box_triangles(lo, hi):
    """12 triangles of the six planes of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    quads = [
        ((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)),  # floor
        ((x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)),  # lid
        ((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)),  # y = y0
        ((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)),  # y = y1
        ((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)),  # x = x0
        ((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)),  # x = x1
    ]
    tris = []
    for a, b, c, d in quads:
        tris.append(a + b + c)
        tris.append(a + c + d)
    return np.array(tris, dtype=np.float64)


def box_lattice(n=10000, dims=None, radius=0.005, density=2000.0, spacing_factor=2.2, jitter=0.1, seed=0,
                spring=1e6, damper=0.1, friction=0.3, velocity=None):
    """Equal spheres on a jittered cubic lattice inside a 6-plane box."""
    if dims is None:
        if n == 10000:
            dims = (25, 20, 20)
        else:
            a = int(math.ceil(n ** (1.0 / 3.0)))
            dims = (a, a, int(math.ceil(n / (a * a))))
    nx, ny, nz = dims
    s = spacing_factor * radius
    rng = np.random.default_rng(seed)
    idx = np.arange(nx * ny * nz)[:n]
    i, j, k = idx // (ny * nz), (idx // nz) % ny, idx % nz
    centers = np.stack([(i + 0.5) * s, (j + 0.5) * s, (k + 0.5) * s], axis=1)
    centers += rng.uniform(-jitter * radius, jitter * radius, size=centers.shape)
    tris = box_triangles((0.0, 0.0, 0.0), (nx * s, ny * s, (nz + 5) * s))
    m = pmx.Model(parnum=n, ellnum=n, trinum=len(tris), pairnum=1)
    _spheres(m, centers, np.full(n, radius), density)
    m.tri[:] = tris.T
    m.tricol[:] = 2
    _granular(m, spring, damper, friction)
    if velocity is not None:
        m.linear[:] = np.asarray(velocity, dtype=np.float64).reshape(3, -1)
    return m


def hopper_triangles(ntheta=100, nz=25, r_out=0.1, height=2.4, angle_deg=60.0):
    """Conical hopper wall (axis z, outlet radius r_out at z = 0, walls at
    angle_deg from horizontal) as 2*ntheta*nz triangles."""
    slope = 1.0 / math.tan(math.radians(angle_deg))
    th = np.linspace(0.0, 2.0 * math.pi, ntheta + 1)
    zz = np.linspace(0.0, height, nz + 1)
    rr = r_out + zz * slope
    tris = []
    for a in range(nz):
        for b in range(ntheta):
            p00 = (rr[a] * math.cos(th[b]), rr[a] * math.sin(th[b]), zz[a])
            p01 = (rr[a] * math.cos(th[b + 1]), rr[a] * math.sin(th[b + 1]), zz[a])
            p10 = (rr[a + 1] * math.cos(th[b]), rr[a + 1] * math.sin(th[b]), zz[a + 1])
            p11 = (rr[a + 1] * math.cos(th[b + 1]), rr[a + 1] * math.sin(th[b + 1]), zz[a + 1])
            tris.append(p00 + p01 + p11)
            tris.append(p00 + p11 + p10)
    return np.array(tris, dtype=np.float64)


def hopper(n=2000000, seed=1, density=2000.0, r_median=0.002, sigma=0.25, r_min=0.001, r_max=0.004,
           ntri=5000, spring=1e6, damper=0.1, friction=0.3, pour_speed=1.0, jitter_speed=0.5, gap=0.0002):
    """Polydisperse spheres poured into a conical hopper.

    Radii ~ lognormal(ln r_median, sigma) clipped to [r_min, r_max].  Centres
    sit on a lattice of spacing 2*r_max + gap, each jittered within its own
    cell, so no two spheres overlap.  Initial velocity (0, 0, -pour_speed)
    plus a uniform +-jitter_speed per component.
    """
    rng = np.random.default_rng(seed)
    radii = np.clip(rng.lognormal(math.log(r_median), sigma, size=n), r_min, r_max)
    s = 2.0 * r_max + gap
    m_side = max(1, int(math.ceil((n / 2.0) ** (1.0 / 3.0))))  # lattice m x m x 2m
    layers = int(math.ceil(n / (m_side * m_side)))
    idx = np.arange(n)
    i, j, k = idx % m_side, (idx // m_side) % m_side, idx // (m_side * m_side)
    half = 0.5 * m_side * s
    slope = 1.0 / math.tan(math.radians(60.0))
    r_out = 0.05 + 0.02 * (m_side * s)
    z0 = max(0.05, (math.sqrt(2.0) * half + 0.02 - r_out) / slope)
    centers = np.stack([(i + 0.5) * s - half, (j + 0.5) * s - half, z0 + (k + 0.5) * s], axis=1)
    room = (0.5 * s - radii)[:, None]
    centers += rng.uniform(-1.0, 1.0, size=centers.shape) * room
    height = z0 + layers * s + 0.05
    nth = max(8, int(round(math.sqrt(ntri / 2.0 * 4.0))))
    nzs = max(1, int(round(ntri / 2.0 / nth)))
    tris = hopper_triangles(ntheta=nth, nz=nzs, r_out=r_out, height=height)
    m = pmx.Model(parnum=n, ellnum=n, trinum=len(tris), pairnum=1)
    _spheres(m, centers, radii, density)
    m.tri[:] = tris.T
    m.tricol[:] = 2
    _granular(m, spring, damper, friction)
    v = rng.uniform(-jitter_speed, jitter_speed, size=(3, n))
    v[2] -= pour_speed
    m.linear[:] = v
    return m
