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


def _cuboid_faces():
    """12 triangles (node numbers of a box's 8 corners, i*4 + j*2 + k for
    offsets (i, j, k) in {0,1}^3), outward oriented."""
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    tris = []
    for a, b, c, d in quads:
        tris.append((a, b, c))
        tris.append((a, c, d))
    return tris


def spring_lattice(nx=40, ny=40, nz=20, size=(0.2, 0.2, 0.4), gap=0.01, density=1800.0, seed=0,
                   stiffness=(1e5, 1e7), curve_step=0.005, damping_ratio=0.05, anchors=True, prescribed_base=False):
    """Lattice of meshed cuboid bodies joined to their 6 neighbours by SPRING
    elements with random 5-point piecewise-linear force-stroke curves.

    Each body is a MESH of 8 nodes and 12 triangles (its mass properties are
    those of a solid cuboid).  Springs run between facing face centres
    (follower direction, initial stroke = gap), curves pass through (0, 0)
    with strokes curve_step*(-2, -1, 0, 1, 2) and slopes uniform in
    `stiffness` (seeded); dashpots use a critical damping ratio.  With
    anchors, every bottom body also hangs on a spring to a fixed point
    `gap` below its bottom face (part2 = -1), so a base excitation can be
    applied in the base frame as a time-dependent gravity.  With
    prescribed_base, the bottom layer instead gets PRESCRIBE entries of kind
    'av' (linear acceleration from base_prescription, angular velocity zero),
    the survey's drive of this configuration.
    """
    rng = np.random.default_rng(seed)
    a, b, c = size
    n = nx * ny * nz
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    idx = lambda i, j, k: (i * ny + j) * nz + k
    centers = np.stack([ii * (a + gap) + 0.5 * a, jj * (b + gap) + 0.5 * b, kk * (c + gap) + 0.5 * c + gap], axis=1)
    mass = density * a * b * c
    m = pmx.Model(parnum=n, ellnum=0, trinum=12 * n, pairnum=1)
    m.position[:3] = centers.T
    m.position[3:] = centers.T
    m.mass[:] = mass
    m.invm[:] = 1.0 / mass
    m.inertia[0] = mass * (b * b + c * c) / 12.0
    m.inertia[4] = mass * (a * a + c * c) / 12.0
    m.inertia[8] = mass * (a * a + b * b) / 12.0
    for k in (0, 4, 8):
        m.inverse[k] = 1.0 / m.inertia[k]
    # mesh: 8 corner nodes, 12 faces per body
    m.set_mesh(8 * n, 12 * n)
    off = np.array([[(q >> 2) & 1, (q >> 1) & 1, q & 1] for q in range(8)], dtype=np.float64) - 0.5
    corners = centers[:, None, :] + off[None, :, :] * np.array([a, b, c])
    corners = corners.reshape(-1, 3)
    m.nodes[:3] = corners.T
    m.nodes[3:] = corners.T
    m.nodpart[:] = np.repeat(np.arange(n, dtype=np.int32), 8)
    faces = np.array(_cuboid_faces(), dtype=np.int32)
    fac = (np.arange(n)[:, None, None] * 8 + faces[None, :, :]).reshape(-1, 3)
    m.facnod[:] = fac.T
    m.factri[:] = np.arange(12 * n, dtype=np.int32)
    m.tri[:] = np.concatenate([corners[fac[:, 0]], corners[fac[:, 1]], corners[fac[:, 2]]], axis=1).T
    m.tricol[:] = 1
    m.triobs[:] = np.repeat(np.arange(n, dtype=np.int32), 12)
    # springs: (part1, part2, point1, point2)
    sp = []
    for d, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        sel = (ii + di < nx) & (jj + dj < ny) & (kk + dk < nz)
        p1 = idx(ii[sel], jj[sel], kk[sel])
        p2 = idx(ii[sel] + di, jj[sel] + dj, kk[sel] + dk)
        half = 0.5 * np.array([a, b, c])[d]
        e = np.zeros(3)
        e[d] = 1.0
        x1 = centers[p1] + half * e
        x2 = centers[p2] - half * e
        sp.append((p1, p2, x1, x2))
    if anchors:
        bot = np.nonzero(kk == 0)[0]
        x1 = centers[bot] - np.array([0.0, 0.0, 0.5 * c])
        x2 = x1 - np.array([0.0, 0.0, gap])
        sp.append((bot, np.full(len(bot), -1), x1, x2))
    p1 = np.concatenate([s[0] for s in sp])
    p2 = np.concatenate([s[1] for s in sp])
    x1 = np.concatenate([s[2] for s in sp])
    x2 = np.concatenate([s[3] for s in sp])
    order = np.lexsort((p2, p1))      # sort_springs: by (part1, part2)
    p1, p2, x1, x2 = p1[order], p2[order], x1[order], x2[order]
    S = len(p1)
    m.set_springs(S, 5 * S, S, 0)
    m.sprtype[:] = pmx.SPRING_NONLINEAR_ELASTIC
    m.sprpart[0] = p1
    m.sprpart[1] = p2
    m.sprpnt[0, :3] = x1.T
    m.sprpnt[0, 3:] = x1.T
    m.sprpnt[1, :3] = x2.T
    m.sprpnt[1, 3:] = x2.T
    slopes = rng.uniform(stiffness[0], stiffness[1], size=(S, 4))
    h = curve_step
    fx = np.array([-2 * h, -h, 0.0, h, 2 * h])
    fy = np.stack([-(slopes[:, 1] + slopes[:, 0]) * h, -slopes[:, 1] * h, np.zeros(S), slopes[:, 2] * h,
                   (slopes[:, 2] + slopes[:, 3]) * h], axis=1)
    m.spring[0] = np.tile(fx, S)
    m.spring[1] = fy.ravel()
    m.spridx[:] = 5 * np.arange(S + 1, dtype=np.int32)
    m.dashpot[0] = damping_ratio
    m.dashpot[1] = damping_ratio
    m.dashidx[:] = np.arange(S + 1, dtype=np.int32)
    m.unidx[:] = 0
    m.sprflg[:] = pmx.SPRDIR_FOLLOWER
    dif = x2 - x1
    m.stroke0[:] = np.sqrt((dif * dif).sum(axis=1))
    if prescribed_base:
        bot = np.nonzero(kk == 0)[0]
        m.set_prescribed(len(bot))
        m.prspart[:] = bot
        m.prskind[0] = 1      # linear: acceleration
        m.prskind[1] = 0      # angular: velocity
        m.prscol[:] = 0       # one history drives the whole base
    return m


def base_prescription(times, amplitude=9.81, frequency=5.0):
    """Per-step table [nsteps][1][6] of one history driving a base layer:
    horizontal acceleration a(t) = amplitude sin(2 pi f t) and zero angular
    velocity (PRESCRIBE kind 'av', input.cpp:4247-4421)."""
    t = np.asarray(times)
    tab = np.zeros((len(t), 1, 6))
    tab[:, 0, 0] = amplitude * np.sin(2.0 * math.pi * frequency * t)
    return tab


def base_excitation(times, amplitude=9.81, frequency=5.0, g=9.81):
    """Gravity table (per step) of a base shaken horizontally with
    a(t) = amplitude*sin(2 pi f t), seen in the base frame."""
    t = np.asarray(times)
    gt = np.zeros((len(t), 3))
    gt[:, 0] = -amplitude * np.sin(2.0 * math.pi * frequency * t)
    gt[:, 2] = -g
    return gt


def sweep_instance(i, nmin=1000, nmax=64000, radius=0.005):
    """Instance i of the acceleration sweep: a box_lattice set-up with a
    seeded particle count in [nmin, nmax], base amplitude in [0.1, 2] g and
    frequency in [1, 20] Hz (seed = instance number).  Returns
    (model, amplitude m/s^2, frequency Hz)."""
    rng = np.random.default_rng(1000003 + i)
    n = int(rng.integers(nmin, nmax + 1))
    amp = float(rng.uniform(0.1, 2.0)) * 9.81
    freq = float(rng.uniform(1.0, 20.0))
    return box_lattice(n=n, seed=i, radius=radius), amp, freq


def dense_instance(i, nmin=1000, nmax=64000, radius=0.005, spacing_factor=1.98, jitter=0.004):
    """Instance i of the sweep (same seeded count, amplitude and frequency)
    as a confined dense pack: the lattice spacing 1.98 r makes every sphere
    overlap its six neighbours by 0.02 r, and the six walls are placed where
    each outer sphere overlaps them by the same 0.02 r, so the pack stays
    compressed (its contact forces, ~100 N, dwarf gravity and the base
    excitation) and every sphere keeps its contacts: about 3 contact pairs
    per sphere, the regime of a settled pack.  The jitter (0.004 r) stays
    below the overlap.  Returns (model, amplitude m/s^2, frequency Hz)."""
    rng = np.random.default_rng(1000003 + i)
    n = int(rng.integers(nmin, nmax + 1))
    amp = float(rng.uniform(0.1, 2.0)) * 9.81
    freq = float(rng.uniform(1.0, 20.0))
    a = int(math.ceil(n ** (1.0 / 3.0)))
    dims = (a, a, int(math.ceil(n / (a * a))))
    m = box_lattice(n=n, dims=dims, seed=i, radius=radius, spacing_factor=spacing_factor, jitter=jitter)
    s = spacing_factor * radius
    inset = 0.5 * s - (radius - (2.0 - spacing_factor) * radius)   # wall overlap = neighbour overlap
    tris = box_triangles((inset, inset, inset), (a * s - inset, a * s - inset, dims[2] * s - inset))
    m.tri[:] = tris.T
    return m, amp, freq


def sweep(instances=512, nmin=1000, nmax=64000):
    """The batched sweep: independent instances of sweep_instance."""
    return [sweep_instance(i, nmin, nmax) for i in range(instances)]


def sweep_forcing(amp, freq, nsteps, step, curtime=0.0):
    """Per-step gravity of a sweep instance: g(t) = (A sin 2 pi f t, 0, -9.81)
    evaluated at each step's curtime, as the reference's GRAVITY callback
    would be (input.cpp:4424-4503)."""
    t = curtime + step * np.arange(nsteps)
    return base_excitation(t, amplitude=-amp, frequency=freq)


def stirred_box(n=2000, seed=0, piston_speed=-0.2, stir_rate=20.0):
    """box_lattice with two moving obstacles: a piston plate over the pack
    descending at piston_speed (m/s) and a vertical blade through the pack
    turning about the box's vertical axis at stir_rate (rad/s).  Returns the
    model and a function giving the obstacle velocity table of dem()
    (obstab [nsteps][2][6], obstab0 [2][6])."""
    m = box_lattice(n=n, seed=seed)
    lo = m.tri.reshape(3, 3, -1).min(axis=(0, 2))
    hi = m.tri.reshape(3, 3, -1).max(axis=(0, 2))
    top = float(m.center[2].max() + 0.006)
    cx, cy = 0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1])
    w = 0.45 * min(hi[0] - lo[0], hi[1] - lo[1])
    piston = [(lo[0] + 1e-3, lo[1] + 1e-3, top, hi[0] - 1e-3, lo[1] + 1e-3, top, hi[0] - 1e-3, hi[1] - 1e-3, top),
              (lo[0] + 1e-3, lo[1] + 1e-3, top, hi[0] - 1e-3, hi[1] - 1e-3, top, lo[0] + 1e-3, hi[1] - 1e-3, top)]
    zb = float(m.center[2].min())
    blade = [(cx - w, cy, zb, cx + w, cy, zb, cx + w, cy, top - 0.01),
             (cx - w, cy, zb, cx + w, cy, top - 0.01, cx - w, cy, top - 0.01)]
    T0 = m.trinum
    extra = np.array(piston + blade, dtype=np.float64)
    m.trinum = T0 + 4
    m.tri = np.ascontiguousarray(np.concatenate([m.tri, extra.T], axis=1))
    m.tricol = np.concatenate([m.tricol, np.full(4, 2, dtype=np.int32)])
    m.triobs = np.concatenate([m.triobs, np.array([-2, -2, -3, -3], dtype=np.int32)])
    m.set_obstacles(2)
    m.trirng[0] = (T0, T0 + 2)
    m.trirng[1] = (T0 + 2, T0 + 4)
    m.obspnt[0] = (cx, cy, top)
    m.obspnt[1] = (cx, cy, 0.5 * (zb + top))

    def table(nsteps):
        row = np.zeros((2, 6))
        row[0, 5] = piston_speed
        row[1, 2] = stir_rate
        return np.tile(row, (nsteps, 1, 1)), row.copy()
    return m, table


def serve_request(i, seed=0, median=20000, sigma=0.8, nmin=2000, nmax=500000, smin=500, smax=20000,
                  spring_share=0.3, step=1e-5):
    """Request i of the serving stream: a box set-up of spheres (70 %) or a
    spring lattice of meshed bodies (30 %), sized lognormal (median, sigma)
    clipped to [nmin, nmax] spheres or bodies, running smin-smax steps under
    the sweep's base excitation (amplitude 0.1-2 g, 1-20 Hz).  Returns
    (model, duration, step, forcing)."""
    rng = np.random.default_rng([seed, i])
    n = int(np.clip(rng.lognormal(math.log(median), sigma), nmin, nmax))
    nsteps = int(rng.integers(smin, smax + 1))
    amp = float(rng.uniform(0.1, 2.0)) * 9.81
    freq = float(rng.uniform(1.0, 20.0))
    if rng.uniform() < spring_share:
        a = max(2, int(round((n / 2.0) ** (1.0 / 3.0))))
        m = spring_lattice(nx=a, ny=a, nz=max(1, n // (a * a)), seed=int(rng.integers(1 << 30)), anchors=False,
                           prescribed_base=True)
        times = step * np.arange(nsteps)
        f = pmx.forcing(nsteps, gravity=(0.0, 0.0, -9.81), prstab=base_prescription(times, amp, freq))
    else:
        m = box_lattice(n=n, seed=int(rng.integers(1 << 30)))
        f = pmx.forcing(nsteps, gtab=sweep_forcing(amp, freq, nsteps, step))
    return m, (nsteps - 0.5) * step, step, f


def serve_stream(count, rate=20.0, seed=0, **kw):
    """Poisson arrivals (rate per second) of serve_request's requests:
    [(arrival time s, model, duration, step, forcing)]."""
    rng = np.random.default_rng(seed + 7919)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    return [(float(t[i]),) + serve_request(i, seed, **kw) for i in range(count)]
