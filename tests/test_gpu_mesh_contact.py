"""GPU parity of sphere contact with meshed particles' triangles against the
oracle: state, contact pairs and forces bit for bit, with the reactions on
the meshed particles summed per triangle (k_react); many contacts on one
triangle (its reaction slots grow and the step repeats); free blocks struck
by spheres; a batch against its instances alone; CRITICAL and the adaptive
step with such contacts."""
import numpy as np
import pytest

from parmec_amd import parmec as P, pmx
from tests.helpers import compare

pytestmark = pytest.mark.gpu

HEX = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def cuboid(x0, size, mat, color):
    nodes = [c for i, j, k in HEX for c in (x0[0] + i * size[0], x0[1] + j * size[1], x0[2] + k * size[2])]
    return P.MESH(nodes, [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, color)


def plate_deck(n=12, seed=0, r=0.04):
    """a 2 x 2 meshed plate held still (prescribed zero velocity) under an
    n x n layer of spheres resting on it, sliding with friction"""
    P.RESET()
    mat = P.MATERIAL(1000.0, 1e6, 0.25)
    plate = cuboid((-1.0, -1.0, -0.2), (2.0, 2.0, 0.2), mat, 2)
    P.PRESCRIBE(plate, lambda t: (0.0, 0.0, 0.0), lambda t: (0.0, 0.0, 0.0))
    rng = np.random.default_rng(seed)
    for i in range(n):
        for j in range(n):
            x = -0.9 + 1.8 * (i + 0.5) / n + rng.uniform(-0.01, 0.01)
            y = -0.9 + 1.8 * (j + 0.5) / n + rng.uniform(-0.01, 0.01)
            s = P.SPHERE((x, y, r * rng.uniform(0.99, 1.01)), r, mat, 1)
            P.VELOCITY(s, (rng.normal() * 0.1, rng.normal() * 0.1, -0.05), (0.0, 0.0, 0.0))
    P.GRANULAR(0, 0, 1e6, 0.9, 0.3)
    P.GRAVITY(0.0, 0.0, -9.81)


def blocks_deck(seed=1):
    """eight free meshed blocks struck by spheres shot at them, no gravity"""
    P.RESET()
    mat = P.MATERIAL(1000.0, 1e6, 0.25)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                cuboid((0.6 * i, 0.6 * j, 0.6 * k), (0.5, 0.5, 0.5), mat, 2)
    rng = np.random.default_rng(seed)
    for q in range(60):
        c = rng.uniform(-0.2, 1.3, 3)
        d = c - np.array([0.55, 0.55, 0.55])
        c = np.array([0.55, 0.55, 0.55]) + d / np.abs(d).max() * 0.64   # 0.04 outside the blocks
        s = P.SPHERE(tuple(float(v) for v in c), 0.05, mat, 1)
        v = -(c - 0.55) / np.linalg.norm(c - 0.55) * 8.0
        P.VELOCITY(s, tuple(float(x) for x in v), tuple(float(x) for x in rng.normal(size=3)))
    P.GRANULAR(0, 0, 1e6, 0.1, 0.2)


def _run_deck(make_engine, deck, script):
    deck()
    e = make_engine()
    P._use_engine(e)
    r = script(e)
    m = P.state().copy()
    c = e.contacts()
    e.close()
    return m, c, r


def _both(deck, script):
    from oracle import orc
    g = _run_deck(lambda: pmx.Engine(critical=1), deck, script)
    o = _run_deck(lambda: orc.Oracle(0), deck, script)
    return g, o


def test_spheres_on_held_plate_match_oracle():
    print("case 144 spheres on a held meshed plate, 600 steps", flush=True)

    def script(e):
        P.DEM(600 * 2e-5 - 1e-5, 2e-5)
        return e.stats()

    (mg, cg, sg), (mc, cc, _) = _both(plate_deck, script)
    tri_contacts = int((cc["slave"] < 0).sum())
    print("contacts", len(cc), "on the plate", tri_contacts, "row capacity", sg["row_capacity"], flush=True)
    assert tri_contacts > 100
    compare(mg, cg, mc, cc)


def test_spheres_strike_free_blocks_match_oracle():
    print("case 60 spheres shot at 8 free meshed blocks, 800 steps", flush=True)

    def script(e):
        P.DEM(800 * 2e-5 - 1e-5, 2e-5)
        return None

    (mg, cg, _), (mc, cc, _) = _both(blocks_deck, script)
    moved = np.abs(mc.linear[:, :8]).max()
    print("block speed", moved, "contacts", len(cc), flush=True)
    assert moved > 0.0
    compare(mg, cg, mc, cc)


def test_mesh_contact_batch_equals_alone():
    print("case batch of 3 plate decks", flush=True)
    models = []
    for k, n in enumerate((6, 9, 12)):
        plate_deck(n=n, seed=k)
        models.append(P.state().copy())
    # the plate's prescribed zero velocity as a table, like the deck builds it
    steps = 300
    out = []
    for m in models:
        ncol = int(m.prscol.max()) + 1
        out.append(pmx.forcing(steps, gravity=(0.0, 0.0, -9.81), prstab=np.zeros((steps, ncol, 6))))
    duration = steps * 2e-5 - 1e-5
    eng = pmx.Engine()
    ms = [m.copy() for m in models]
    eng.upload_batch(ms)
    eng.dem_batch(duration, 2e-5, out)
    eng.download_batch(ms)
    cs = [eng.contacts(i) for i in range(len(ms))]
    eng.close()
    for i, m in enumerate(models):
        e = pmx.Engine()
        ma = m.copy()
        e.upload(ma)
        e.dem(duration, 2e-5, out[i])
        e.download(ma)
        ca = e.contacts()
        e.close()
        assert int((ca["slave"] < 0).sum()) > 0
        compare(ms[i], cs[i], ma, ca)


def test_critical_and_adaptive_with_mesh_contact_match_oracle():
    print("case plate deck: DEM, CRITICAL, adaptive DEM", flush=True)

    def script(e):
        P.DEM(200 * 2e-5 - 1e-5, 2e-5)
        per = P.CRITICAL(perparticle=65)   # the plate and 64 spheres
        P.DEM(0.0005, 2e-5, adaptive=0.5)   # ~2,800 steps: the rotational estimate of 4 cm spheres is ~4e-7 s
        return per

    (mg, cg, pg), (mc, cc, pc) = _both(lambda: plate_deck(n=8), script)
    assert len(pg) == len(pc)
    for a, b in zip(pg, pc):
        assert a[1] == b[1] and (a[0] == b[0] or (a[0] != a[0] and b[0] != b[0])), (a, b)
    compare(mg, cg, mc, cc)
