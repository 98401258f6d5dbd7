"""The benchmark workloads at their full sizes:

- the 512-instance sweep (1,000-64,000 spheres each, 16.9M spheres), stirred
  so that contacts form at once, batched in mixed precision for 200 steps:
  sampled instances (the smallest, a middle one, the largest) bitwise equal
  to their runs alone; and the same batch in FP64 for 100 steps, sampled
  instances bitwise equal to the oracle;
- the serving stream at its stated distribution (box set-ups and spring
  lattices of 2,000-500,000 bodies, lognormal around 20,000, 500-20,000
  steps): every request served, the two smallest bitwise equal to their runs
  alone.
"""
import numpy as np
import pytest

from parmec_amd import pmx, workloads
from tests.helpers import compare

pytestmark = pytest.mark.gpu


def _sweep(nsteps, step, count=512):
    models, forcings = [], []
    for i in range(count):
        m, amp, freq = workloads.sweep_instance(i, 1000, 64000)
        m.linear[:] = np.random.default_rng(i).normal(scale=0.3, size=m.linear.shape)
        models.append(m)
        forcings.append(pmx.forcing(nsteps, gtab=workloads.sweep_forcing(amp, freq, nsteps, step)))
    return models, forcings


def _batch(models, forcings, duration, step, prec, sample):
    eng = pmx.Engine(precision=prec)
    ms = [m.copy() for m in models]
    eng.upload_batch(ms)
    eng.dem_batch(duration, step, forcings)
    st = eng.stats()
    eng.download_batch(ms)
    out = {i: (ms[i], eng.contacts(i)) for i in sample}
    eng.close()
    return out, st


def test_sweep512_mixed_sample_equals_alone():
    step, nsteps = 1e-5, 200
    print(f"case sweep 512 instances, mixed, {nsteps} steps, sample vs alone", flush=True)
    models, forcings = _sweep(nsteps, step)
    sizes = np.array([m.parnum for m in models])
    sample = [int(sizes.argmin()), 256, int(sizes.argmax())]
    duration = (nsteps - 0.5) * step
    got, st = _batch(models, forcings, duration, step, pmx.PMX_MIXED, sample)
    print("spheres", int(sizes.sum()), "pairs", st["pairs"], "rebuilds", st["rebuilds"], flush=True)
    for i in sample:
        e = pmx.Engine(precision=pmx.PMX_MIXED)
        ma = models[i].copy()
        e.upload(ma)
        e.dem(duration, step, forcings[i])
        e.download(ma)
        ca = e.contacts()
        e.close()
        print("instance", i, "spheres", models[i].parnum, "contacts", len(ca), flush=True)
        assert len(ca) > 0
        compare(got[i][0], got[i][1], ma, ca)


def test_sweep512_fp64_sample_equals_oracle():
    from oracle import orc
    step, nsteps = 1e-5, 100
    print(f"case sweep 512 instances, fp64, {nsteps} steps, sample vs oracle", flush=True)
    models, forcings = _sweep(nsteps, step)
    sample = [3, 300]
    duration = (nsteps - 0.5) * step
    got, _ = _batch(models, forcings, duration, step, pmx.PMX_FP64, sample)
    for i in sample:
        o = orc.Oracle(0)
        mc = models[i].copy()
        o.upload(mc)
        o.dem(duration, step, forcings[i])
        o.download(mc)
        cc = o.contacts()
        o.close()
        print("instance", i, "spheres", models[i].parnum, "contacts", len(cc), flush=True)
        compare(got[i][0], got[i][1], mc, cc)


def test_serving_stream_at_its_sizes():
    count = 10
    print(f"case serving stream: {count} requests of the stated distribution", flush=True)
    stream = workloads.serve_stream(count, rate=50.0, seed=3)
    sizes = [m.parnum for (_, m, _, _, _) in stream]
    steps = [int(round(d / h + 0.5)) for (_, _, d, h, _) in stream]
    print("bodies", sizes, "steps", steps, flush=True)
    assert max(sizes) > 20000 and max(steps) > 5000
    srv = pmx.Server(precision=pmx.PMX_MIXED, gather_s=0.2)
    ids = [srv.submit(m, d, h, f) for (_, m, d, h, f) in stream]
    outs = []
    for rid, (_, m, d, h, f) in zip(ids, stream):
        mm = m.copy()
        info = srv.wait(rid, mm, timeout=600.0)
        assert info["status"] == pmx.REQ_DONE
        print("request", rid, "bodies", m.parnum, "batch", info["batch"], flush=True)
        outs.append(mm)
    srv.close()
    for k in np.argsort(sizes)[:2]:
        _, m, d, h, f = stream[k]
        e = pmx.Engine(precision=pmx.PMX_MIXED)
        ma = m.copy()
        e.upload(ma)
        e.dem(d, h, f)
        e.download(ma)
        e.close()
        for name in ("position", "linear", "angular", "rotation", "force", "torque"):
            assert np.array_equal(getattr(outs[k], name), getattr(ma, name)), (k, name)
