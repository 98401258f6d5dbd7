"""The resident request loop: every request's result is bitwise the result
of running it alone (one upload + one dem call), whether it ran alone, in a
batch with others or beside batches admitted at other times; the HBM budget
decides admission."""
import time

import numpy as np
import pytest

from parmec_amd import pmx, workloads
from tests.helpers import spring_modes_model

pytestmark = pytest.mark.gpu

OUT = ("position", "linear", "angular", "rotation", "center", "force", "torque", "tri", "nodes", "stroke", "sprfrc")


def _requests(step=1e-5):
    reqs = []
    for i in range(4):
        m, amp, freq = workloads.sweep_instance(i, 500, 3000)
        m.linear[:] = np.random.default_rng(i).normal(scale=0.5, size=m.linear.shape)
        n = 150 + 90 * i
        reqs.append((m, (n - 0.5) * step, step, pmx.forcing(n, gtab=workloads.sweep_forcing(amp, freq, n, step))))
    s = spring_modes_model(seed=3)
    reqs.append((s, 220.5 * step, step, pmx.forcing(gravity=(0.0, 0.0, -9.81))))
    e = workloads.spring_lattice(nx=3, ny=3, nz=2, seed=4)
    reqs.append((e, 260.5 * step, step, pmx.forcing(gravity=(0.0, 0.0, -9.81))))
    pl = workloads.spring_lattice(nx=4, ny=3, nz=3, seed=6, anchors=False, prescribed_base=True)
    reqs.append((pl, 240.5 * step, step, pmx.forcing(241, gravity=(0.0, 0.0, -9.81),
                                                      prstab=workloads.base_prescription(step * np.arange(241)))))
    b, table = workloads.stirred_box(n=1200, seed=5)
    tab, tab0 = table(300)
    reqs.append((b, 299.5 * step, step, pmx.forcing(300, gravity=(0.0, 0.0, -9.81), obstab=tab, obstab0=tab0)))
    return reqs


def _alone(prec, m, duration, step, f):
    e = pmx.Engine(precision=prec)
    mm = m.copy()
    e.upload(mm)
    e.dem(duration, step, f)
    e.download(mm)
    e.close()
    return mm


def _serve(prec, reqs, gap=0.0, **kw):
    srv = pmx.Server(precision=prec, **kw)
    ids = []
    for (m, d, h, f) in reqs:
        ids.append(srv.submit(m, d, h, f))
        if gap:
            time.sleep(gap)
    outs, infos = [], []
    for rid, (m, d, h, f) in zip(ids, reqs):
        mm = m.copy()
        infos.append(srv.wait(rid, mm, timeout=90.0))
        print("request", rid, "done in batch of", infos[-1]["batch"], flush=True)
        outs.append(mm)
    srv.close()
    return outs, infos


def _same(a, b):
    for name in OUT:
        x, y = getattr(a, name, None), getattr(b, name, None)
        if x is None:
            continue
        assert np.array_equal(x, y), name


@pytest.mark.parametrize("prec", [pmx.PMX_FP64, pmx.PMX_MIXED])
def test_served_equals_alone(prec):
    reqs = _requests()
    print(f"case serve precision {prec}: {len(reqs)} requests, batched and staggered", flush=True)
    together, info1 = _serve(prec, reqs, gather_s=0.5)           # one batch per kind of request
    staggered, info2 = _serve(prec, reqs, gap=0.03, gather_s=0.0)  # separate, overlapping batches
    longest, info3 = _serve(prec, reqs, gap=0.01, gather_s=0.1, max_batch=3, longest_first=True)
    # a batch holds one kind: spheres, or bodies joined by springs or meshes
    kinds = [int(m.ellnum > 0) + 2 * int(m.sprnum > 0 or m.facnum > 0) for (m, d, h, f) in reqs]
    assert [i["batch"] for i in info1] == [kinds.count(k) for k in kinds]
    assert all(i["status"] == pmx.REQ_DONE for i in info1 + info2 + info3)
    for k, (m, d, h, f) in enumerate(reqs):
        alone = _alone(prec, m, d, h, f)
        _same(together[k], alone)
        _same(staggered[k], alone)
        _same(longest[k], alone)
        assert together[k].curtime == alone.curtime


@pytest.mark.parametrize("prec", [pmx.PMX_FP64, pmx.PMX_MIXED])
def test_continuous_batching_equals_alone(prec):
    """Requests submitted one by one join the running batch of their kind
    (continuous batching: the batch pauses after a group of steps, carries
    its unfinished requests over with their contact rows, tree snapshots and
    time, and resumes with the new ones): each result is still bitwise its
    solo run, and the requests ran in batches of changing size."""
    reqs = _requests()
    print(f"case continuous batching precision {prec}: {len(reqs)} requests, staggered", flush=True)
    outs, info = _serve(prec, reqs, gap=0.004, gather_s=0.0, repack_s=1e-4)
    assert all(i["status"] == pmx.REQ_DONE for i in info)
    print("batch sizes", [i["batch"] for i in info], flush=True)
    assert max(i["batch"] for i in info) > 1
    bad = []
    for k, (m, d, h, f) in enumerate(reqs):
        alone = _alone(prec, m, d, h, f)
        for name in OUT:
            x, y = getattr(outs[k], name, None), getattr(alone, name, None)
            if x is not None and not np.array_equal(x, y):
                bad.append((k, name, float(np.abs(np.asarray(x) - np.asarray(y)).max())))
        if outs[k].curtime != alone.curtime or outs[k].stepnum != alone.stepnum:
            bad.append((k, "time", outs[k].curtime - alone.curtime, outs[k].stepnum, alone.stepnum))
    assert not bad, bad


def test_budget_serialises_admission():
    reqs = [r for r in _requests()[:3]]
    _, info = _serve(pmx.PMX_FP64, reqs[:1])
    budget = int(1.05 * max(_serve(pmx.PMX_FP64, [r])[1][0]["bytes"] for r in reqs))
    print("case budget", budget, flush=True)
    _, info = _serve(pmx.PMX_FP64, reqs, hbm_budget=budget)
    for a, b in zip(info, info[1:]):
        assert b["admit_s"] >= a["done_s"] - 1e-3
    with pytest.raises(pmx.PmxError):
        _serve(pmx.PMX_FP64, reqs[:1], hbm_budget=1024)
