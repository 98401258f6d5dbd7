"""Time histories (HISTORY, input.cpp:3157-3555; sampling parmec.cpp:2386-2394,
2509-2514; values output.cpp:5143-5520) sampled on the device inside one
dem call.

Pinned two ways: against a closed form (a free sphere moving and spinning
at constant rates: the leapfrog puts the positions 1.5 steps past each
sample's time, the spin is exact), and against the oracle's states on a
spring lattice (no contacts, so splitting a dem call into one-step calls
changes nothing), with output_history's formulas restated in numpy below.
"""
import math

import numpy as np
import pytest

from parmec_amd import parmec as P, pmx, workloads

pytestmark = pytest.mark.gpu


def test_history_closed_form():
    """A sphere (centre (0.5, 0.5, 0.5)) at v = (1, 0, 0), spin (0, 0, 1):
    like the reference's tests/history.py (a unit cube there)."""
    print("case history closed form", flush=True)
    P.RESET()
    mat = P.MATERIAL(1e3, 1e9, 0.25)
    p = P.SPHERE((0.5, 0.5, 0.5), 0.5, mat, 1)
    P.VELOCITY(p, linear=(1.0, 0.0, 0.0), angular=(0.0, 0.0, 1.0))
    t = P.HISTORY("TIME")
    dx0 = P.HISTORY("DX", p)
    dx1 = P.HISTORY("DX", p, (1.0, 1.0, 1.0))
    dy1 = P.HISTORY("DY", p, (1.0, 1.0, 1.0))
    vx = P.HISTORY("VX", p)
    oz = P.HISTORY("OZ", p)
    vl1 = P.HISTORY("|V|", p, (1.0, 1.0, 1.0))
    h = 0.1
    P.DEM(1.0, h)
    P.RESET()
    # the sample before the first step, then one per step from the second
    # (curtime k h crosses curtime_history + h at k = 1, 2, ...; 0.1 added
    # ten times stays below 1.0, so the call takes 11 steps)
    times = pmx.step_times(1.0, h, 0.0)
    n = len(t)
    assert n == len(times) == 11 and t[0] == 0.0
    assert dx0[0] == 0.0 and dx1[0] == 0.0 and dy1[0] == 0.0
    assert abs(vx[0] - 1.0) < 1e-15 and abs(oz[0] - 1.0) < 1e-15
    for k in range(1, n):
        assert t[k] == times[k]
        s = (k + 1.5) * h                   # half step at the start, then k + 1 steps
        assert abs(dx0[k] - s) < 1e-12
        assert abs(vx[k] - 1.0) < 1e-15 and abs(oz[k] - 1.0) < 1e-12
        # body point (1, 1, 1): centre + R(s) (0.5, 0.5, 0.5), R a rotation by s about z
        c, sn = math.cos(s), math.sin(s)
        px = 0.5 + s + 0.5 * c - 0.5 * sn
        py = 0.5 + 0.5 * sn + 0.5 * c
        assert abs(dx1[k] - (px - 1.0)) < 1e-12
        assert abs(dy1[k] - (py - 1.0)) < 1e-12
        # the point velocity as output.cpp:5203-5208 forms it, v + a x o with
        # a = p - x = (0.5 c - 0.5 s, 0.5 s + 0.5 c, 0.5): (1 + a_y, -a_x, 0)
        ax, ay = 0.5 * c - 0.5 * sn, 0.5 * sn + 0.5 * c
        assert abs(vl1[k] - math.hypot(1.0 + ay, ax)) < 1e-12


def _his_values(m, ent, lst, pt):
    """output_history (output.cpp:5143-5520) on a downloaded state."""
    N, S = m.parnum, m.sprnum
    pos, lin, ang, rot = m.position, m.linear, m.angular, m.rotation
    frc, trq = m.force, m.torque

    def L3(a, b, c):
        return math.sqrt(a * a + b * b + c * c)
    if pt is not None:
        k = lst[0]
        x = pos[0:3, k]
        X = pos[3:6, k]
        v = lin[:, k]
        o = ang[3:6, k]
        R = rot[:, k]
        Q = [pt[0] - X[0], pt[1] - X[1], pt[2] - X[2]]
        p = [x[r] + (R[r] * Q[0] + R[r + 3] * Q[1] + R[r + 6] * Q[2]) for r in range(3)]
        a = [p[r] - x[r] for r in range(3)]
        w = [v[0] + a[1] * o[2] - a[2] * o[1], v[1] + a[2] * o[0] - a[0] * o[2], v[2] + a[0] * o[1] - a[1] * o[0]]
        table = {"PX": p[0], "PY": p[1], "PZ": p[2], "DX": p[0] - pt[0], "DY": p[1] - pt[1], "DZ": p[2] - pt[2],
                 "VX": w[0], "VY": w[1], "VZ": w[2], "|V|": L3(*w), "OZ": o[2], "FX": frc[0, k], "TZ": trq[2, k]}
        return table[ent]
    value = 0.0
    for k in lst:
        if ent == "PX":
            value += pos[0, k]
        elif ent == "DZ":
            value += pos[2, k] - pos[5, k]
        elif ent == "|V|":
            value += L3(lin[0, k], lin[1, k], lin[2, k])
        elif ent == "OY":
            value += ang[4, k]
        elif ent == "FZ":
            value += frc[2, k]
        elif ent == "|T|":
            value += L3(trq[0, k], trq[1, k], trq[2, k])
        elif ent == "LENGTH":
            sp = m.sprpnt.reshape(2, 6, S)
            value += L3(sp[1, 0, k] - sp[0, 0, k], sp[1, 1, k] - sp[0, 1, k], sp[1, 2, k] - sp[0, 2, k])
        elif ent == "STROKE":
            value += m.stroke.reshape(3, S)[0, k]
        elif ent == "F":
            value += m.sprfrc.reshape(3, S)[0, k]
        elif ent == "SF":
            value += m.sprfrc.reshape(3, S)[1, k]
        else:
            raise AssertionError(ent)
    return value / float(len(lst))


def test_history_matches_oracle_states():
    """Spring lattice under gravity: particle-list, point and spring
    histories sampled at every step end on the device in one call, against
    the same quantities of the oracle's state stepped one step per call."""
    from oracle import orc
    print("case history spring lattice 60 steps", flush=True)
    m0 = workloads.spring_lattice(nx=3, ny=3, nz=3, seed=4, anchors=False)
    nsteps, h = 60, 1e-4
    f = pmx.forcing(nsteps, gravity=(0.0, 0.0, -9.81))
    S = m0.sprnum
    hs = [("PX", [0, 4, 13], None), ("DZ", [26], None), ("|V|", [1, 2, 3, 5], None), ("OY", [7], None),
          ("FZ", [8, 9], None), ("|T|", [10], None), ("LENGTH", [0, 1, S - 1], None), ("STROKE", [2], None),
          ("F", [3, 4], None), ("SF", [5], None),
          ("PX", [11], (0.01, 0.02, 0.03)), ("VY", [12], (0.0, 0.1, -0.05)), ("|V|", [13], (0.05, 0.05, 0.05)),
          ("FX", [14], (0.0, 0.0, 0.0)), ("TZ", [15], (0.1, 0.0, 0.0))]
    g = pmx.Engine()
    mg = m0.copy()
    g.upload(mg)
    g.set_history([e for e, _, _ in hs], [l for _, l, _ in hs], [p for _, _, p in hs])
    pre = g.history_sample()
    hisrow = np.arange(nsteps, dtype=np.int32)
    hisout = np.zeros((nsteps, len(hs)))
    fh = pmx.forcing(nsteps, gravity=(0.0, 0.0, -9.81), hisrow=hisrow, hisout=hisout)
    g.dem((nsteps - 0.5) * h, h, fh)
    g.download(mg)
    g.close()
    c = orc.Oracle(0)
    mc = m0.copy()
    c.upload(mc)
    for i, (e, l, p) in enumerate(hs):
        assert pre[i] == _his_values(mc, e, l, p), ("before the first step", e)
    f1 = pmx.forcing(1, gravity=(0.0, 0.0, -9.81))
    for k in range(nsteps):
        c.dem(0.5 * h, h, f1)
        c.download(mc)
        for i, (e, l, p) in enumerate(hs):
            want = _his_values(mc, e, l, p)
            assert hisout[k, i] == want, (k, e, l, p, hisout[k, i], want)
    c.close()
    assert np.array_equal(mg.position, mc.position)
