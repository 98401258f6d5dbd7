"""Batches: every instance of a batch must give bitwise what it gives alone
(state, contact pairs and tangential history), whatever else is in the batch.
The single runs are themselves pinned to the oracle by the parity tests."""
import numpy as np
import pytest

from parmec_amd import pmx, workloads
from helpers import compare, spring_modes_model

pytestmark = pytest.mark.gpu


def _alone(m, duration, step, f):
    eng = pmx.Engine()
    mm = m.copy()
    eng.upload(mm)
    eng.dem(duration, step, f)
    eng.download(mm)
    out = mm, eng.contacts(), eng.stats()
    eng.close()
    return out


def _batch(models, duration, step, forcings, calls=1):
    eng = pmx.Engine()
    ms = [m.copy() for m in models]
    eng.upload_batch(ms)
    for _ in range(calls):
        eng.dem_batch(duration, step, forcings)
    eng.download_batch(ms)
    cs = [eng.contacts(i) for i in range(len(ms))]
    st = eng.stats()
    eng.close()
    return ms, cs, st


def _sweep_models(count, nmin, nmax, nsteps, step, stir=0.5):
    """Sweep instances, stirred with seeded random velocities (m/s) so that
    contacts form and break within a short run."""
    models, forcings = [], []
    for i in range(count):
        m, amp, freq = workloads.sweep_instance(i, nmin, nmax)
        m.linear[:] = np.random.default_rng(i).normal(scale=stir, size=m.linear.shape)
        models.append(m)
        forcings.append(pmx.forcing(nsteps, gtab=workloads.sweep_forcing(amp, freq, nsteps, step)))
    return models, forcings


def test_batch_equals_alone_sweep():
    """Sweep instances of uneven size, each with its own gravity table."""
    step, nsteps = 1e-5, 300
    print(f"case sweep batch 8 instances x {nsteps} steps", flush=True)
    models, forcings = _sweep_models(8, 300, 3000, nsteps, step)
    duration = (nsteps - 0.5) * step
    ms, cs, _ = _batch(models, duration, step, forcings)
    for i, m in enumerate(models):
        ma, ca, _ = _alone(m, duration, step, forcings[i])
        assert len(ca) > 0
        compare(ms[i], cs[i], ma, ca)


def test_batch_mixed_kinds_and_pairings():
    """Spheres with different pairing tables, a spring model and a repartition
    in one instance, stepped in one batch over two dem calls."""
    step = 1e-5
    a = workloads.box_lattice(n=1500, seed=5)
    b = workloads.box_lattice(n=900, seed=6, friction=0.0, damper=0.3)
    c = workloads.box_lattice(n=4000, seed=3)
    xc = c.center[:3].mean(axis=1, keepdims=True)
    c.linear[:] = -75.0 * (c.center[:3] - xc)       # converging flow: leaf overflow
    d = spring_modes_model(seed=2)
    e = workloads.spring_lattice(nx=4, ny=3, nz=3, seed=1)
    models = [a, b, c, d, e]
    fs = [pmx.forcing(gravity=(0.0, 0.0, -9.81)), pmx.forcing(gravity=(1.0, 0.0, -9.81)), pmx.forcing(),
          pmx.forcing(gravity=(0.0, 0.0, -9.81), damping=(0.1, 0.1, 0.1, 0.0, 0.0, 0.0)),
          pmx.forcing(gravity=(0.0, 0.0, -9.81))]
    duration = 100 * step - 0.5 * step
    print("case mixed batch (spheres, pairings, springs, repartition), 2 calls", flush=True)
    ms, cs, st = _batch(models, duration, step, fs, calls=2)
    assert st["rebuilds"] > 2
    for i, m in enumerate(models):
        eng = pmx.Engine()
        mm = m.copy()
        eng.upload(mm)
        for _ in range(2):
            eng.dem(duration, step, fs[i])
        eng.download(mm)
        ca = eng.contacts()
        eng.close()
        compare(ms[i], cs[i], mm, ca)
        if m.sprnum:
            for name in ("stroke", "sprfrc", "sprpnt", "sprdir"):
                assert np.array_equal(getattr(ms[i], name), getattr(mm, name)), name


def test_batch_order_invariance():
    """The same instances in another order give the same per-instance results."""
    step, nsteps = 1e-5, 200
    models, forcings = _sweep_models(5, 500, 2500, nsteps, step)
    duration = (nsteps - 0.5) * step
    print("case batch order invariance", flush=True)
    ms1, cs1, _ = _batch(models, duration, step, forcings)
    perm = [3, 0, 4, 2, 1]
    ms2, cs2, _ = _batch([models[p] for p in perm], duration, step, [forcings[p] for p in perm])
    for k, p in enumerate(perm):
        compare(ms2[k], cs2[k], ms1[p], cs1[p])


def test_partial_rebuild_batch():
    """One instance of a batch overflows its leaves again and again: each
    rebuild re-sorts and re-partitions only that instance (the others keep
    their trees), and every instance still gives bitwise its solo result."""
    step, nsteps = 1e-5, 150
    models, forcings = _sweep_models(5, 500, 3000, nsteps, step)
    c = workloads.box_lattice(n=4000, seed=3)
    xc = c.center[:3].mean(axis=1, keepdims=True)
    c.linear[:] = -75.0 * (c.center[:3] - xc)       # converging flow: leaf overflow
    models.insert(2, c)
    forcings.insert(2, pmx.forcing(nsteps))
    duration = (nsteps - 0.5) * step
    print("case partial rebuilds in a batch of 6", flush=True)
    ms, cs, st = _batch(models, duration, step, forcings, calls=2)
    assert st["partial_rebuilds"] > 2
    assert st["partial_rebuilds"] == st["rebuilds"] - 1   # all but the build at the call's start
    for i, m in enumerate(models):
        eng = pmx.Engine()
        mm = m.copy()
        eng.upload(mm)
        for _ in range(2):
            eng.dem(duration, step, forcings[i])
        eng.download(mm)
        ca = eng.contacts()
        eng.close()
        compare(ms[i], cs[i], mm, ca)


def test_coincident_overflow_batch():
    """Two identical converging instances of 20,000 spheres (many blocks of
    the binning launch each) overflow their leaves in the same steps, beside
    a neutral one.  The binning that overflows flags every overflowing
    instance, so one re-partition per such step serves both, as each would
    re-partition alone; every instance is bitwise its solo run."""
    step, nsteps = 1e-5, 150
    c = workloads.box_lattice(n=20000, seed=3)
    xc = c.center[:3].mean(axis=1, keepdims=True)
    c.linear[:] = -75.0 * (c.center[:3] - xc)       # converging flow: leaf overflows
    neutral = workloads.box_lattice(n=3000, seed=4)
    models = [c, neutral, c.copy()]
    fs = [pmx.forcing(), pmx.forcing(gravity=(0.0, 0.0, -9.81)), pmx.forcing()]
    duration = (nsteps - 0.5) * step
    print("case coincident leaf overflows: 2 x 20000 converging + 3000 at rest", flush=True)
    ms, cs, st = _batch(models, duration, step, fs)
    alone = [_alone(m, duration, step, fs[i]) for i, m in enumerate(models)]
    print("rebuilds batch", st["rebuilds"], "alone", [a[2]["rebuilds"] for a in alone], flush=True)
    assert alone[0][2]["rebuilds"] > 2 and alone[1][2]["rebuilds"] == 1
    assert st["rebuilds"] == alone[0][2]["rebuilds"]
    for i in range(3):
        compare(ms[i], cs[i], alone[i][0], alone[i][1])


def test_batch_pause_changes_nothing():
    """A batch paused after a step (the first forcing's output point, where a
    caller may read the state or start a clock) continues bitwise as without
    the pause; the callback sees the step once."""
    step, nsteps = 1e-5, 120
    print(f"case batch pause, 4 instances x {nsteps} steps", flush=True)
    models, forcings = _sweep_models(4, 300, 2000, nsteps, step)
    duration = (nsteps - 0.5) * step
    ms, cs, _ = _batch(models, duration, step, forcings)
    seen = []
    m0, a0, f0 = workloads.sweep_instance(0, 300, 2000)
    paused = list(forcings)
    paused[0] = pmx.forcing(nsteps, gtab=workloads.sweep_forcing(a0, f0, nsteps, step), outsteps=[59],
                            on_output=seen.append)
    mp, cp, _ = _batch(models, duration, step, paused)
    assert seen == [59]
    for i in range(len(models)):
        compare(mp[i], cp[i], ms[i], cs[i])


def test_device_window_changes_nothing():
    """pmx_set_window only records events: the batch steps bitwise as
    without it, and the window's device time is positive and below the
    call's wall time."""
    import ctypes as C
    import time
    step, nsteps = 1e-5, 80
    print(f"case device window, 4 instances x {nsteps} steps", flush=True)
    models, forcings = _sweep_models(4, 300, 2000, nsteps, step)
    duration = (nsteps - 0.5) * step
    ms, cs, _ = _batch(models, duration, step, forcings)
    eng = pmx.Engine()
    mw = [m.copy() for m in models]
    eng.upload_batch(mw)
    eng.lib.pmx_set_window.argtypes = [C.c_void_p, C.c_int, C.c_int]
    eng.lib.pmx_set_window(eng.h, 30, 69)
    t0 = time.perf_counter()
    eng.dem_batch(duration, step, forcings)
    wall = time.perf_counter() - t0
    st = eng.stats()
    eng.download_batch(mw)
    cw = [eng.contacts(i) for i in range(len(mw))]
    eng.dem_batch(duration, step, forcings)
    assert eng.stats()["window_ms"] == -1.0      # one call only
    eng.close()
    assert 0.0 < st["window_ms"] < 1e3 * wall
    for i in range(len(models)):
        compare(mw[i], cw[i], ms[i], cs[i])
