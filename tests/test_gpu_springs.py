"""GPU parity of SPRING elements and meshed particles against the oracle:
identical FP64 state bit for bit (positions, velocities, rotations, forces,
torques, spring points, directions, strokes, forces, yield limits, flags)."""
import numpy as np
import pytest

from parmec_amd import pmx, workloads
from tests.helpers import compare, run, spring_modes_model

pytestmark = pytest.mark.gpu

SPRING_STATE = ("sprpnt", "sprdir", "stroke", "sprfrc", "yield_", "sprsdsp", "sprflg", "nodes", "tri")


def _both(model, duration, step, forcing):
    from oracle import orc
    g, c = pmx.Engine(), orc.Oracle(0)
    a = run(g, model, duration, step, forcing)
    b = run(c, model, duration, step, forcing)
    g.close(); c.close()
    return a, b


def _compare_springs(mg, mc):
    for name in SPRING_STATE:
        a, b = getattr(mg, name), getattr(mc, name)
        assert np.array_equal(a, b), f"{name}: max diff {float(np.abs(a - b).max()) if a.size else 0.0}"


def test_spring_modes():
    print("case spring modes", flush=True)
    m = spring_modes_model(seed=1)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(m, 500 * 1e-4 - 0.5e-4, 1e-4, f)
    assert np.isfinite(mc.position).all()
    compare(mg, cg, mc, cc)
    _compare_springs(mg, mc)
    assert (mc.sprflg & pmx.SPRING_YIELDED).any()


@pytest.mark.parametrize("dims,steps", [((6, 5, 4), 500), ((40, 40, 20), 200)])
def test_spring_lattice(dims, steps):
    print(f"case spring lattice {dims} steps={steps}", flush=True)
    m = workloads.spring_lattice(*dims)
    times = pmx.step_times(steps * 1e-5 - 0.5e-5, 1e-5)
    f = pmx.forcing(len(times), gtab=workloads.base_excitation(times))
    (mg, cg, sg), (mc, cc, sc) = _both(m, steps * 1e-5 - 0.5e-5, 1e-5, f)
    assert sg["steps"] == steps
    compare(mg, cg, mc, cc)
    _compare_springs(mg, mc)


@pytest.mark.parametrize("dims,steps,calls", [((6, 5, 4), 400, 2), ((40, 40, 20), 150, 1)])
def test_spring_lattice_prescribed_base(dims, steps, calls):
    """The spring-lattice configuration driven as proposed: bottom layer under
    PRESCRIBE kind 'av' (horizontal base acceleration 1 g at 5 Hz, zero
    angular velocity), gravity on; over one or two dem calls."""
    from oracle import orc
    print(f"case prescribed-base lattice {dims} steps={steps} calls={calls}", flush=True)
    m = workloads.spring_lattice(*dims, anchors=False, prescribed_base=True, seed=2)
    out = []
    for eng in (pmx.Engine(), orc.Oracle(0)):
        mm = m.copy()
        eng.upload(mm)
        for call in range(calls):
            times = pmx.step_times(steps * 1e-5 - 0.5e-5, 1e-5, mm.curtime)
            f = pmx.forcing(len(times), gravity=(0.0, 0.0, -9.81),
                            prstab=workloads.base_prescription(times))
            eng.dem(steps * 1e-5 - 0.5e-5, 1e-5, f)
            eng.download(mm)
        out.append((mm, eng.contacts()))
        eng.close()
    (mg, cg), (mc, cc) = out
    compare(mg, cg, mc, cc)
    _compare_springs(mg, mc)
    bot = mc.prspart
    assert np.abs(mc.linear[0, bot]).max() > 0.0      # the base moved
    np.testing.assert_array_equal(mc.angular[3:, bot], 0.0)


def test_deck_mesh_lattice():
    """A spring lattice built entirely through the deck API (MESH bodies,
    SPRING curves, PRESCRIBE base drive, GRAVITY) runs on the device
    bitwise as on the oracle, nodes and surface triangles included."""
    from oracle import orc
    from parmec_amd import parmec as P
    from tests.test_deck_mesh import deck_lattice
    print("case deck-built lattice 6x5x4 (MESH + SPRING + PRESCRIBE), 2 calls", flush=True)
    out = []
    for make in (lambda: pmx.Engine(), lambda: orc.Oracle(0)):
        P.RESET()
        deck_lattice(6, 5, 4)
        P.GRAVITY(0.0, 0.0, -9.81)
        P._use_engine(make())
        P.DEM(200 * 1e-5 - 0.5e-5, 1e-5)
        P.DEM(150 * 1e-5 - 0.5e-5, 1e-5)
        out.append(P.state().copy())
        P.RESET()
    g, c = out
    for name in ("position", "linear", "angular", "rotation", "force", "torque", "stroke", "sprfrc", "sprpnt",
                 "sprdir", "nodes", "tri"):
        a, b = getattr(g, name), getattr(c, name)
        assert np.array_equal(a, b), (name, float(np.abs(a - b).max()))
    assert np.abs(g.sprfrc).max() > 0.0
