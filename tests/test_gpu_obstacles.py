"""Moving obstacles on the device: bitwise parity with the oracle (triangle
motion, the per-step reach of moving triangles, the obstacle branch of the
contact law) and batch invariance."""
import numpy as np
import pytest

from parmec_amd import pmx, workloads
from tests.helpers import compare, run

pytestmark = pytest.mark.gpu


def test_stirred_box_parity():
    from oracle import orc
    print("case stirred box 2000 (piston + blade), 600 steps", flush=True)
    m, table = workloads.stirred_box(n=2000, seed=1)
    nsteps, h = 600, 1e-5
    tab, tab0 = table(nsteps)
    f = pmx.forcing(nsteps, gravity=(0.0, 0.0, -9.81), obstab=tab, obstab0=tab0)
    g = pmx.Engine()
    mg, cg, sg = run(g, m, (nsteps - 0.5) * h, h, f)
    c = orc.Oracle(0)
    mc, cc, sc = run(c, m, (nsteps - 0.5) * h, h, f)
    g.close(); c.close()
    tri_contacts = int((cc["slave"] < 0).sum())
    moving = int(np.isin(-cc["slave"] - 1, np.arange(m.trinum - 4, m.trinum)).sum())
    print("contacts", len(cc), "triangle", tri_contacts, "moving", moving, flush=True)
    assert moving > 0
    compare(mg, cg, mc, cc)
    assert np.array_equal(mg.tri, mc.tri)
    assert np.array_equal(mg.obspnt, mc.obspnt)


def test_stirred_box_two_calls_and_batch():
    """Two dem calls (obstacle velocities persist across calls) and the same
    instance inside a batch beside a plain one."""
    print("case stirred box batch, 2 calls", flush=True)
    m, table = workloads.stirred_box(n=1500, seed=2)
    plain = workloads.box_lattice(n=800, seed=4)
    nsteps, h = 250, 1e-5
    tab, tab0 = table(nsteps)
    f = pmx.forcing(nsteps, gravity=(0.0, 0.0, -9.81), obstab=tab, obstab0=tab0)
    fp = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    e = pmx.Engine()
    ma, ca, _ = run(e, m, (nsteps - 0.5) * h, h, f, calls=2)
    e.close()
    e = pmx.Engine()
    ms = [plain.copy(), m.copy()]
    e.upload_batch(ms)
    for _ in range(2):
        e.dem_batch((nsteps - 0.5) * h, h, [fp, f])
    e.download_batch(ms)
    cb = e.contacts(1)
    e.close()
    compare(ms[1], cb, ma, ca)
    assert np.array_equal(ms[1].tri, ma.tri)
