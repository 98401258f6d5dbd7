"""GPU parity: the HIP step path against the CPU oracle on the same inputs.

Bar: identical contact-pair sets and bit-identical FP64 state (positions,
velocities, rotations, centres, forces, torques, tangential contact state);
the stated tolerance of the path is 1e-10 relative, bitwise equality is what
the two implementations deliver with FP contraction off on both sides.
"""
import numpy as np
import pytest

from parmec_amd import pmx, workloads
from tests.helpers import compare, run

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engines():
    from oracle import orc
    return pmx.Engine(), orc.Oracle(0)


def _both(engines, model, duration, step, forcing=None, calls=1):
    gpu, cpu = engines
    g = pmx.Engine()
    c = type(cpu)(0)
    mg, cg, sg = run(g, model, duration, step, forcing, calls)
    mc, cc, sc = run(c, model, duration, step, forcing, calls)
    g.close(); c.close()
    return (mg, cg, sg), (mc, cc, sc)


def test_two_spheres_head_on(engines):
    m = workloads.box_lattice(n=2, dims=(2, 1, 1), jitter=0.0)
    m.linear[0, 0] = 0.5
    m.linear[0, 1] = -0.5
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 2e-3, 1e-5, pmx.forcing())
    compare(mg, cg, mc, cc)


@pytest.mark.parametrize("n,steps", [(500, 300), (4000, 200)])
def test_box_lattice_gravity(engines, n, steps):
    print(f"case box_lattice n={n} steps={steps}", flush=True)
    m = workloads.box_lattice(n=n, seed=3, spacing_factor=2.05, jitter=0.02)
    rng = np.random.default_rng(7)
    m.linear[:] = rng.uniform(-1.0, 1.0, size=m.linear.shape)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, steps * 1e-5, 1e-5, f)
    assert sg["steps"] == sc["steps"]
    assert len(cc) > 0
    compare(mg, cg, mc, cc)


@pytest.mark.parametrize("damping", [(0.5, 0.5, 0.5, 0.2, 0.2, 0.2), (0.0, 0.0, 0.0, 0.3, 0.3, 0.3)])
def test_box_lattice_damped(engines, damping):
    """Linear and angular damping (dynamics.ispc:64-91) on spheres in and out
    of contact: the damped force and torque are the stored ones, the
    undamped part of the pair is formed from gravity where a row is empty."""
    print(f"case box_lattice damped {damping}", flush=True)
    m = workloads.box_lattice(n=2000, seed=4, spacing_factor=2.05, jitter=0.02)
    rng = np.random.default_rng(8)
    m.linear[:] = rng.uniform(-1.0, 1.0, size=m.linear.shape)
    m.angular[:3] = rng.uniform(-20.0, 20.0, size=(3, m.parnum))
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81), damping=damping)
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 250 * 1e-5, 1e-5, f, calls=2)
    assert sg["steps"] == sc["steps"]
    assert len(cc) > 0
    compare(mg, cg, mc, cc)


def test_hopper_small(engines):
    print("case hopper n=20000 steps=200", flush=True)
    m = workloads.hopper(n=20000, seed=5, ntri=800, jitter_speed=5.0, gap=0.00002)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 200 * 2e-6, 2e-6, f)
    assert len(cc) > 50
    assert sg["leaf_slots"] == 96   # the slotted leaf-record layout
    compare(mg, cg, mc, cc)


def test_box10k_full_run(engines):
    """10,000 equal spheres (r = 5 mm, 2000 kg/m^3) on a 25x20x20 jittered
    lattice (spacing 2.2r, jitter +-0.1r) in a 6-plane box, kn = 1e6,
    damping ratio 0.1, mu = 0.3, gravity -9.81 z, dt = 1e-5, 2,000 steps."""
    print("case box10k steps=2000", flush=True)
    m = workloads.box_lattice(n=10000, seed=0)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 2000 * 1e-5 - 0.5e-5, 1e-5, f)
    assert sg["steps"] == sc["steps"] == 2000
    assert len(cc) > 20
    compare(mg, cg, mc, cc)


def test_hopper_2m_short(engines):
    """The 2,000,000-sphere hopper at full size, 6 steps (the oracle's cost
    bounds the length)."""
    print("case hopper n=2000000 steps=6", flush=True)
    m = workloads.hopper(n=2000000, seed=1)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 6 * 2e-6 - 1e-6, 2e-6, f)
    assert sg["steps"] == 6
    compare(mg, cg, mc, cc)


@pytest.mark.parametrize("mode", [1, 2])
def test_profiled_paths_match_oracle(engines, mode):
    """Profiling mode 1 replays, every 8th step, a copy of the captured step
    graph with event record nodes around each kernel (fresh events per
    launch), the plain graph otherwise; mode 2 enqueues every step eagerly
    with an event pair per launch.  Both must give the oracle's state and
    time the detection launches of the steps they sample."""
    import ctypes as C
    from oracle import orc
    m = workloads.box_lattice(n=4000, seed=3, spacing_factor=2.05, jitter=0.02)
    rng = np.random.default_rng(7)
    m.linear[:] = rng.uniform(-1.0, 1.0, size=m.linear.shape)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    g = pmx.Engine()
    assert g.lib.pmx_set_profiling(g.h, mode) == 0
    mg, cg, sg = run(g, m, 150 * 1e-5 - 0.5e-5, 1e-5, f, calls=2)
    arr = (pmx.pmx_kernel_time * 32)()
    cnt = C.c_int(0)
    g.lib.pmx_get_profile.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    assert g.lib.pmx_get_profile(g.h, C.cast(arr, C.c_void_p), 32, C.byref(cnt)) == 0
    prof = {arr[k].name.decode(): arr[k] for k in range(cnt.value)}
    g.close()
    # both calls' sampled steps, less any step aborted by a leaf overflow (not counted)
    sampled = 2 * 150 if mode == 2 else 2 * len(range(0, 150, 8))
    assert sampled - 2 * sg["rebuilds"] <= prof["detect"].launches <= sampled
    assert prof["detect"].ms > 0.0
    assert sg["graph_timed"] == (1 if mode == 1 else 0)
    c = orc.Oracle(0)
    mc, cc, sc = run(c, m, 150 * 1e-5 - 0.5e-5, 1e-5, f, calls=2)
    c.close()
    compare(mg, cg, mc, cc)


def test_leaf_overflow_repartition(engines):
    """Spheres converge on the centre so inner leaves exceed 96 entries:
    the tree is rebuilt from the current centres and the step repeated
    (parmec.cpp:2420-2427), identically on both sides."""
    print("case converging flow (re-partitioning)", flush=True)
    m = workloads.box_lattice(n=4000, seed=3)
    xc = m.center[:3].mean(axis=1, keepdims=True)
    m.linear[:] = -75.0 * (m.center[:3] - xc)
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 200 * 1e-5 - 0.5e-5, 1e-5, pmx.forcing())
    assert sc["rebuilds"] > 5
    assert sg["rebuilds"] == sc["rebuilds"]
    compare(mg, cg, mc, cc)


def test_full_leaves_with_contacts(engines):
    """Leaves between 65 and 96 spheres while contacts form: a lattice at
    2.2 radii spacing contracts by 15 % over 100 steps (density x 1.6), so
    built leaves of up to 64 fill past 64 before they overflow, and
    neighbours touch from step ~61.  Pins the detection paths taken by
    leaves of more than 64 spheres (second reach-mask slot, the generic
    own-leaf pair loop) against the oracle."""
    print("case converging flow with contacts (leaves over 64)", flush=True)
    m = workloads.box_lattice(n=4000, seed=5)
    xc = m.center[:3].mean(axis=1, keepdims=True)
    m.linear[:] = -150.0 * (m.center[:3] - xc)
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 100 * 1e-5 - 0.5e-5, 1e-5, pmx.forcing())
    assert len(cc) > 1000
    assert sc["rebuilds"] >= 2   # the first build, then at least one leaf past 96
    assert sg["rebuilds"] == sc["rebuilds"]
    compare(mg, cg, mc, cc)


def test_separate_binning(engines, monkeypatch):
    """Binning at every step's start (k_bin; PMX_FUSED_BIN=0) instead of at
    the end of the previous step's dynamics gives the oracle's result, with
    the same re-partitions, through leaf overflows (the fused form's
    overflow completes its step and aborts the next one)."""
    print("case converging flow, separate binning", flush=True)
    monkeypatch.setenv("PMX_FUSED_BIN", "0")
    m = workloads.box_lattice(n=4000, seed=3)
    xc = m.center[:3].mean(axis=1, keepdims=True)
    m.linear[:] = -75.0 * (m.center[:3] - xc)
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 200 * 1e-5 - 0.5e-5, 1e-5, pmx.forcing())
    assert sc["rebuilds"] > 5
    assert sg["rebuilds"] == sc["rebuilds"]
    compare(mg, cg, mc, cc)


def test_csr_neighbour_lists(engines, monkeypatch):
    """Neighbour lists in the CSR layout (used when the padded layout would
    not fit) give the oracle's result as the padded layout does."""
    print("case hopper n=20000 steps=100 with CSR neighbour lists", flush=True)
    monkeypatch.setenv("PMX_NEIGHBOUR_LISTS", "csr")
    m = workloads.hopper(n=20000, seed=6, ntri=800, jitter_speed=5.0, gap=0.00002)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 100 * 2e-6, 2e-6, f)
    assert len(cc) > 50
    compare(mg, cg, mc, cc)


def test_packed_leaf_records(engines, monkeypatch):
    """Leaf records packed at the scanned leaf counts (used when the slotted
    layout would take more than 4 slots per sphere) give the oracle's result
    as the slotted layout does."""
    print("case hopper n=20000 steps=100 with packed leaf records", flush=True)
    monkeypatch.setenv("PMX_LEAF_RECORDS", "packed")
    m = workloads.hopper(n=20000, seed=7, ntri=800, jitter_speed=5.0, gap=0.00002)
    f = pmx.forcing(gravity=(0.0, 0.0, -9.81))
    (mg, cg, sg), (mc, cc, sc) = _both(engines, m, 100 * 2e-6, 2e-6, f)
    assert len(cc) > 50
    assert sg["leaf_slots"] == 0
    compare(mg, cg, mc, cc)


@pytest.mark.parametrize("b_x,tri_y,c_y,expect", [
    (5.05, None, 4.5, set()),          # overlapping pair whose boxes never reach each other's leaves
    (5.0, None, 4.5, {(0, 1)}),        # lower x bound exactly on the root's split coordinate: found
    (5.05, 4.05, 4.5, set()),          # triangle box not widened by the radius: contact missed
    (5.05, 4.1, 4.55, {(2, -1)}),      # triangle box on the split coordinate: found
])
def test_q10_hand_built(b_x, tri_y, c_y, expect):
    """The reference's frozen-tree rule on cases whose pair sets are derived
    by hand (tests/test_oracle_tree.py::q10_model): the device finds exactly
    the pairs the rule finds, including the overlapping ones it misses."""
    from tests.test_oracle_tree import q10_model, q10_contacts
    print(f"case q10 b_x={b_x} tri_y={tri_y}", flush=True)
    m = q10_model(b_x=b_x, tri_y=tri_y, c_y=c_y)
    g = pmx.Engine()
    got = q10_contacts(g, m)
    g.close()
    assert got == expect
