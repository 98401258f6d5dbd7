"""Host-side tests: the Parmec input API mirror, the model layout and the
C ABI of libpmx.so (no device calls)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from parmec_amd import parmec as P
from parmec_amd import pmx, workloads
from tests.conftest import ROOT, gpu_available


@pytest.fixture(autouse=True)
def fresh():
    P._S.reset()
    yield
    P._S.reset()


def _deck_box(n=200):
    m = workloads.box_lattice(n=n, seed=4)
    mat = P.MATERIAL(2000.0, 1e6, 0.25)
    for i in range(n):
        P.SPHERE(tuple(m.center[:3, i]), 0.005, mat, 1)
    P.OBSTACLE([tuple(t) for t in m.tri.T], 2)
    P.GRANULAR(0, 0, 1e6, 0.1, 0.3)
    P.GRAVITY(0.0, 0.0, -9.81)
    return m


def test_deck_builds_the_same_soa_as_the_generator():
    ref = _deck_box()
    m = P._build_model()
    for name in ("position", "center", "radii", "mass", "inertia", "tri", "tricol", "triobs", "part", "ellcol",
                 "iparam", "pairs", "ikind", "rotation"):
        np.testing.assert_array_equal(getattr(m, name), getattr(ref, name), err_msg=name)
    assert m.mass[0] == (4.0 / 3.0) * math.pi * 0.005 * 0.005 * 0.005 * 2000.0


def test_argument_errors_follow_the_reference():
    mat = P.MATERIAL(1000.0, 1e6, 0.2)
    with pytest.raises(ValueError):
        P.MATERIAL(-1.0, 1e6, 0.2)
    with pytest.raises(ValueError):
        P.SPHERE((0.0, 0.0, 0.0), -0.1, mat, 1)
    with pytest.raises(ValueError):
        P.SPHERE((0.0, 0.0, 0.0), 0.1, mat + 1, 1)
    with pytest.raises(ValueError):
        P.SPHERE((0.0, 0.0, 0.0), 0.1, mat, 0)
    with pytest.raises(TypeError):
        P.SPHERE((0.0, 0.0), 0.1, mat, 1)
    with pytest.raises(ValueError):
        P.GRANULAR(0, 1, 1e6)
    with pytest.raises(ValueError):
        P.GRAVITY(0, 0, -10)          # ints are TSERIES numbers (input.cpp:2857-2867)
    ts = P.TSERIES([0.0, 0.0, 1.0, -10.0])
    P.GRAVITY(0.0, 0.0, ts)
    with pytest.raises(ValueError):
        P.DEM(-1.0, 1e-3)
    with pytest.raises(ValueError):
        P.MESH([], [], mat, 1)           # no element (the reference would allocate a negative node count)


def test_tseries_interpolation_and_clamping():
    ts = P._S
    k = P.TSERIES([[0.0, 1.0], [1.0, 3.0], [3.0, -1.0]])
    f = ts.tms[k]
    assert f(-1.0) == 1.0 and f(5.0) == -1.0
    assert f(0.5) == 2.0 and f(2.0) == 1.0
    assert ts.tms[P.TSERIES(4.5)](10.0) == 4.5


def test_dem_on_injected_engine_matches_direct_engine():
    from oracle import orc
    _deck_box(100)
    P.VELOCITY(3, (0.5, 0.0, 0.0), (0.0, 0.0, 2.0))
    P._use_engine(orc.Oracle(1))
    P.DEM(50 * 1e-5 - 0.5e-5, 1e-5)
    a = P.state().copy()
    ref = workloads.box_lattice(n=100, seed=4)
    ref.linear[:, 3] = (0.5, 0.0, 0.0)
    ref.angular[3:, 3] = (0.0, 0.0, 2.0)
    ref.angular[:3, 3] = (0.0, 0.0, 2.0)
    o = orc.Oracle(1)
    o.upload(ref)
    o.dem(50 * 1e-5 - 0.5e-5, 1e-5, pmx.forcing(gravity=(0.0, 0.0, -9.81)))
    o.download(ref)
    for name in ("position", "linear", "angular", "rotation"):
        np.testing.assert_array_equal(getattr(a, name), getattr(ref, name), err_msg=name)
    # a second DEM call continues from curtime (parmec.cpp:2386-2409)
    P.DEM(10 * 1e-5 - 0.5e-5, 1e-5)
    assert abs(P.state().curtime - 60e-5) < 1e-15


def test_gravity_callback_is_tabulated_at_step_times():
    from oracle import orc
    mat = P.MATERIAL(1000.0, 1e6, 0.2)
    P.SPHERE((0.0, 0.0, 0.0), 0.1, mat, 1)
    seen = []
    P.GRAVITY(0.0, 0.0, lambda t: (seen.append(t), -9.81 * (1.0 + t))[1])
    P._use_engine(orc.Oracle(1))
    P.DEM(10 * 1e-3 - 0.5e-3, 1e-3)
    assert len(seen) == 10
    np.testing.assert_allclose(seen, np.arange(10) * 1e-3, atol=1e-15)


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "pmx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(pmx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(pmx.LIBPATH)
    syms = _declared_symbols()
    assert len(syms) >= 11
    for s in syms:
        assert hasattr(lib, s), s


def test_step_times_reproduce_the_reference_accumulation():
    # parmec.cpp:2418: time += 0.5*(h+h) while time < duration
    t = pmx.step_times(0.01, 1e-4)
    assert len(t) == 101
    assert len(pmx.step_times((500 - 0.5) * 2e-6, 2e-6)) == 500


@pytest.mark.skipif(gpu_available(), reason="checks the failure path on a host without a device")
def test_engine_fails_loudly_without_a_device():
    with pytest.raises(pmx.PmxError):
        pmx.Engine()


@pytest.mark.parametrize("shape", [{"det_waves": 3}, {"rows_block": 96}, {"det_waves": 16, "rows_block": 128}])
def test_engine_rejects_unknown_block_shapes(shape):
    """Fixed block shapes are checked before any device is opened."""
    with pytest.raises(pmx.PmxError) as err:
        pmx.Engine(tune=1, **shape)
    assert "must be" in str(err.value)


def test_history_registration_errors():
    """HISTORY validates like input.cpp:3157-3555 (no device needed)."""
    from parmec_amd import parmec as P
    P.RESET()
    with pytest.raises(ValueError):
        P.HISTORY("TIME")                      # no particle defined yet
    mat = P.MATERIAL(1e3, 1e9, 0.25)
    p = P.SPHERE((0.0, 0.0, 0.0), 0.5, mat, 1)
    assert P.HISTORY("TIME") == []
    with pytest.raises(ValueError):
        P.HISTORY("NOPE", p)
    with pytest.raises(ValueError):
        P.HISTORY("PX", p + 1)
    with pytest.raises(ValueError):
        P.HISTORY("LENGTH", 0)                 # no spring
    with pytest.raises(ValueError):
        P.HISTORY("PX", p, (1.0, 2.0))
    with pytest.raises(NotImplementedError):
        P.HISTORY("PX", (0.0, 0.0, 0.0, 1.0))
    with pytest.raises(NotImplementedError):
        P.HISTORY("PX", p, h5file="x.h5")
    P.RESET()


def test_history_sampling_rule():
    """parmec.cpp:2509-2514: a step is sampled when its curtime reaches
    curtime_history + interval, and the history time then advances by the
    interval; both sums accumulate in floating point as the reference's do,
    which decides the borderline steps."""
    from parmec_amd import parmec as P, pmx
    h = 0.1
    times = pmx.step_times(1.0, h, 0.0)            # 11 steps: 0.1 added ten times stays below 1.0
    assert len(times) == 11
    rows, ts, ch = P._history_rows(times, 0.0, lambda t: h)
    assert list(rows) == [-1] + list(range(10))    # every step from the second (same sums on both sides)
    assert ts == list(times[1:]) and ch == ts[-1]
    # an interval of two steps: at step 6 curtime is 0.6 but the history time
    # plus 0.2 is 0.6000000000000001, so the sample moves to step 7 (a count
    # of steps would give 2, 4, 6, 8, 10)
    rows, ts, ch = P._history_rows(times, 0.0, lambda t: 2 * h)
    assert [k for k in range(11) if rows[k] >= 0] == [2, 4, 7, 9]
    # an interval that changes with time is evaluated at each step's curtime
    rows, ts, ch = P._history_rows(times, 0.0, lambda t: 0.3 if t < 0.5 else 0.1)
    assert [k for k in range(11) if rows[k] >= 0] == [3, 5, 6, 7, 8, 9, 10]
    # a second call continues from the first call's history time
    _, _, ch0 = P._history_rows(times, 0.0, lambda t: h)
    rows2, _, _ = P._history_rows([1.0, 1.1], ch0, lambda t: h)
    assert list(rows2) == [-1, 0]                  # 1.0 < 0.9999999999999999 + 0.1
