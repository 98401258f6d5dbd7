"""Autotuner: the block shapes it picks (at the first call and again once the
context has taken 256 steps) never change a result.  A batch above the
tuner's 50,000-sphere threshold, run over two calls that cross the second
pass, must give bitwise the state, contact pairs and tangential history of
the same run with fixed shapes (two shapes unlike the defaults)."""
import numpy as np
import pytest

from parmec_amd import pmx, workloads
from helpers import compare

pytestmark = pytest.mark.gpu


def _run(models, forcings, duration, step, precision, calls, **shape):
    eng = pmx.Engine(precision=precision, **shape)
    ms = [m.copy() for m in models]
    eng.upload_batch(ms)
    for _ in range(calls):
        eng.dem_batch(duration, step, forcings)
    eng.download_batch(ms)
    cs = [eng.contacts(i) for i in range(len(ms))]
    st = eng.stats()
    eng.close()
    return ms, cs, st


@pytest.mark.parametrize("precision", [pmx.PMX_FP64, pmx.PMX_MIXED])
def test_tuned_equals_fixed(precision):
    step, nsteps, calls = 1e-5, 200, 2
    print(f"case tune precision {precision}: 4 instances, {calls} calls x {nsteps} steps", flush=True)
    models, forcings = [], []
    for i in range(4):
        m, amp, freq = workloads.sweep_instance(i, 14000, 16000)
        m.linear[:] = np.random.default_rng(i).normal(scale=0.5, size=m.linear.shape)
        models.append(m)
        forcings.append(pmx.forcing(nsteps, gtab=workloads.sweep_forcing(amp, freq, nsteps, step)))
    assert sum(m.ellnum for m in models) >= 50000
    duration = (nsteps - 0.5) * step
    mt, ct, st = _run(models, forcings, duration, step, precision, calls, tune=0)
    print(f"tuned shapes: detect {st['det_waves']} waves, rows {st['rows_block']} threads", flush=True)
    for nw, tpb in ((2, 128), (8, 64)):
        mf, cf, sf = _run(models, forcings, duration, step, precision, calls, tune=1, det_waves=nw, rows_block=tpb)
        assert (sf["det_waves"], sf["rows_block"]) == (nw, tpb)
        for i in range(len(models)):
            assert len(ct[i]) > 0
            compare(mt[i], ct[i], mf[i], cf[i])
