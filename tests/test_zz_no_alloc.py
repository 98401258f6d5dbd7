"""pmx_history_sample_noalloc (no device allocation; its sample passes
through a static per-GPU buffer) against pmx_history_sample on the same
resident state: bitwise equal, on two calls in a row, the second with fewer
histories; more than PMX_HIS_NOALLOC_MAX histories is a clear error."""
import ctypes as C

import numpy as np
import pytest

from parmec_amd import pmx, workloads

pytestmark = pytest.mark.gpu

ENTS = ["PX", "PY", "PZ", "|P|", "VX", "VZ", "|V|", "OX", "|O|", "FX", "FZ", "|F|", "TY", "|T|", "DZ"]


def _noalloc(eng, n):
    out = np.full(max(n, 1), np.nan)
    f = eng.lib.pmx_history_sample_noalloc
    f.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    rc = f(eng.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out[:n]


def test_noalloc_sample_equals_sample():
    print("case history sample, static scratch vs library buffer", flush=True)
    m = workloads.box_lattice(n=1500, seed=3, spacing_factor=2.05, jitter=0.02)
    m.linear[:] = np.random.default_rng(7).uniform(-1.0, 1.0, size=m.linear.shape)
    eng = pmx.Engine()
    mm = m.copy()
    eng.upload(mm)
    eng.dem(200 * 1e-5 - 0.5e-5, 1e-5, pmx.forcing(gravity=(0.0, 0.0, -9.81)))
    rng = np.random.default_rng(1)
    lists = [sorted(rng.choice(m.parnum, size=int(rng.integers(1, 40)), replace=False).tolist()) for _ in ENTS]
    for count in (len(ENTS), 4):   # the second call samples fewer histories
        eng.set_history(ENTS[:count], lists[:count])
        ref = eng.history_sample()
        rc, got = _noalloc(eng, count)
        assert rc == 0, eng.lib.pmx_last_error()
        assert len(got) == count and np.all(np.isfinite(got))
        assert np.array_equal(got, ref), (got, ref)
    eng.dem(50 * 1e-5 - 0.5e-5, 1e-5, pmx.forcing(gravity=(0.0, 0.0, -9.81)))   # a dem call still runs after it
    assert np.array_equal(_noalloc(eng, 4)[1], eng.history_sample())
    eng.close()


def test_noalloc_sample_rejects_too_many():
    eng = pmx.Engine()
    m = workloads.box_lattice(n=100, seed=1)
    mm = m.copy()
    eng.upload(mm)
    eng.set_history(["PX"] * 4097, [[0]] * 4097)
    rc, _ = _noalloc(eng, 0)
    assert rc == -4   # PMX_ECAPACITY
    assert b"at most 4096" in eng.lib.pmx_last_error()
    eng.close()
