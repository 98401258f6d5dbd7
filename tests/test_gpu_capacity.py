"""Contact storage grows instead of failing.  The reference appends contact
chunks without bound (condet.ispc:52-81); the device rows, hits, reverse hits
and per-leaf moving-triangle lists have capacities, and a step that overflows
one is aborted before it changes any state, the capacity grows and the step
is repeated.  Results must be the oracle's (whose lists are unbounded) and
those of a run that had the capacity from the start."""
import math

import numpy as np
import pytest

from parmec_amd import pmx, workloads
from tests.helpers import compare, run

pytestmark = pytest.mark.gpu


def big_among_small(count=70, big=0.004, small=0.001, overlap=2e-5, seed=0):
    """One sphere of radius `big` at the centre of a box, `count` spheres of
    radius `small` on a Fibonacci shell around it, each overlapping it by
    `overlap` and moving slowly inwards: the big sphere holds `count`
    contacts from the first step (more than a 32-entry row)."""
    rng = np.random.default_rng(seed)
    k = np.arange(count) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / count)
    theta = math.pi * (1.0 + 5.0 ** 0.5) * k
    u = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)
    rs = big + small - overlap
    centers = np.concatenate([np.zeros((1, 3)), rs * u]) + 0.01
    radii = np.concatenate([[big], np.full(count, small)])
    tris = workloads.box_triangles((0.0, 0.0, 0.0), (0.02, 0.02, 0.02))
    n = count + 1
    m = pmx.Model(parnum=n, ellnum=n, trinum=len(tris), pairnum=1)
    workloads._spheres(m, centers, radii, 2000.0)
    m.tri[:] = tris.T
    m.tricol[:] = 2
    workloads._granular(m, 1e5, 0.1, 0.3)
    m.linear[:, 1:] = (-0.02 * u + rng.normal(scale=0.005, size=u.shape)).T
    return m


def _two_calls(engine, m, h, f, first=20, second=180):
    """A first dem call of `first` steps (the contacts of the big sphere
    last ~30 steps), then one of `second` steps; the stats of the first."""
    mm = m.copy()
    engine.upload(mm)
    engine.dem((first - 0.5) * h, h, f)
    s1 = engine.stats()
    engine.dem((second - 0.5) * h, h, f)
    engine.download(mm)
    return mm, engine.contacts(), s1


@pytest.mark.parametrize("precision", [pmx.PMX_FP64, pmx.PMX_MIXED])
def test_big_sphere_among_small(precision):
    """70 contacts on one sphere: rows and hits grow (32 -> 128) in the
    first step; FP64 is bitwise the oracle, both precisions bitwise a run
    with 256-entry rows from the start."""
    from oracle import orc
    print(f"case big sphere among 70 small, precision {precision}", flush=True)
    m = big_among_small()
    h = 1e-6
    f = pmx.forcing()
    g = pmx.Engine(precision=precision)
    mg, cg, sg = _two_calls(g, m, h, f)
    g.close()
    print("first call", sg, flush=True)
    assert sg["grown"] >= 1 and sg["row_capacity"] >= 128, sg
    assert sg["max_row"] >= 60, sg
    wide = pmx.Engine(precision=precision, row_capacity=256)
    mw, cw, sw = _two_calls(wide, m, h, f)
    wide.close()
    assert sw["grown"] == 0 and sw["max_row"] == sg["max_row"]
    compare(mg, cg, mw, cw)
    if precision == pmx.PMX_FP64:
        c = orc.Oracle(0)
        mc, cc, _ = _two_calls(c, m, h, f)
        c.close()
        compare(mg, cg, mc, cc)


def test_capacity_growth_in_batch():
    """The big-sphere instance inside a batch: the capacity grows for the
    whole batch, and every instance is bitwise its solo run."""
    print("case big sphere in a batch of 3", flush=True)
    h, nsteps = 1e-6, 150
    models = [workloads.box_lattice(n=600, seed=1), big_among_small(seed=2), workloads.box_lattice(n=900, seed=2)]
    for k, mm in enumerate(models):
        if k != 1:
            mm.linear[:] = np.random.default_rng(k).normal(scale=0.5, size=mm.linear.shape)
    fs = [pmx.forcing(gravity=(0.0, 0.0, -9.81))] * 3
    e = pmx.Engine()
    ms = [mm.copy() for mm in models]
    e.upload_batch(ms)
    e.dem_batch((nsteps - 0.5) * h, h, fs)
    e.download_batch(ms)
    cs = [e.contacts(i) for i in range(3)]
    st = e.stats()
    e.close()
    assert st["grown"] >= 1
    for i, mm in enumerate(models):
        a = pmx.Engine()
        ma, ca, _ = run(a, mm, (nsteps - 0.5) * h, h, fs[i])
        a.close()
        compare(ms[i], cs[i], ma, ca)


def test_moving_triangle_lists_grow():
    """A piston meshed into 800 triangles descends onto a pack: more than 32
    moving triangles reach a leaf, the per-leaf lists grow, and the run is
    bitwise the oracle's."""
    from oracle import orc
    print("case finely meshed piston (800 moving triangles)", flush=True)
    m = workloads.box_lattice(n=1200, seed=5)
    lo = m.tri.reshape(3, 3, -1).min(axis=(0, 2))
    hi = m.tri.reshape(3, 3, -1).max(axis=(0, 2))
    top = float(m.center[2].max() + 0.0052)
    nx = 20
    xs = np.linspace(lo[0] + 1e-3, hi[0] - 1e-3, nx + 1)
    ys = np.linspace(lo[1] + 1e-3, hi[1] - 1e-3, nx + 1)
    tris = []
    for i in range(nx):
        for j in range(nx):
            a = (xs[i], ys[j], top); b = (xs[i + 1], ys[j], top)
            c = (xs[i + 1], ys[j + 1], top); d = (xs[i], ys[j + 1], top)
            tris.append(a + b + c)
            tris.append(a + c + d)
    extra = np.array(tris, dtype=np.float64)
    T0 = m.trinum
    m.trinum = T0 + len(extra)
    m.tri = np.ascontiguousarray(np.concatenate([m.tri, extra.T], axis=1))
    m.tricol = np.concatenate([m.tricol, np.full(len(extra), 2, dtype=np.int32)])
    m.triobs = np.concatenate([m.triobs, np.full(len(extra), -2, dtype=np.int32)])
    m.set_obstacles(1)
    m.trirng[0] = (T0, T0 + len(extra))
    m.obspnt[0] = (0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), top)
    nsteps, h = 300, 1e-5
    row = np.zeros((1, 6))
    row[0, 5] = -0.2
    tab, tab0 = np.tile(row, (nsteps, 1, 1)), row.copy()
    f = pmx.forcing(nsteps, gravity=(0.0, 0.0, -9.81), obstab=tab, obstab0=tab0)
    g = pmx.Engine()
    mg, cg, sg = run(g, m, (nsteps - 0.5) * h, h, f)
    g.close()
    c = orc.Oracle(0)
    mc, cc, _ = run(c, m, (nsteps - 0.5) * h, h, f)
    c.close()
    assert sg["grown"] >= 1, sg
    moving = int((cc["slave"] <= -(T0 + 1)).sum())
    print("contacts", len(cc), "with the piston", moving, flush=True)
    assert moving > 0
    compare(mg, cg, mc, cc)
    assert np.array_equal(mg.tri, mc.tri)
