"""Shared helpers of the parity tests: run the same model through the HIP
engine and the CPU oracle and compare."""
import numpy as np

from parmec_amd import pmx

STATE = ("position", "linear", "angular", "rotation", "center", "force", "torque")


def run(engine, model, duration, step, forcing=None, calls=1):
    m = model.copy()
    engine.upload(m)
    for _ in range(calls):
        engine.dem(duration, step, forcing if forcing is not None else pmx.forcing())
    engine.download(m)
    return m, engine.contacts(), engine.stats()


def max_rel(a, b):
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    s[s == 0] = 1.0
    return float((d / s).max()) if d.size else 0.0


def compare(ma, ca, mb, cb, exact=True, rtol=1e-10):
    """Pair sets must be identical; state bitwise (exact) or within rtol."""
    assert len(ca) == len(cb), f"pair count {len(ca)} != {len(cb)}"
    np.testing.assert_array_equal(ca["master"], cb["master"])
    np.testing.assert_array_equal(ca["slave"], cb["slave"])
    for name in STATE:
        a, b = getattr(ma, name), getattr(mb, name)
        if exact:
            if not np.array_equal(a, b):
                raise AssertionError(f"{name}: max rel diff {max_rel(a, b):.3e}")
        else:
            assert max_rel(a, b) <= rtol, f"{name}: {max_rel(a, b):.3e}"
    if exact:
        np.testing.assert_array_equal(ca["state"], cb["state"])
    else:
        assert max_rel(ca["state"], cb["state"]) <= max(rtol, 1e-8)


def spring_modes_model(seed=0, n=24):
    """Spheres joined by springs of every kind the step path supports, built
    through the deck API (SPHERE, SPRING): follower, constant and planar
    directions, point-to-plane, critical-ratio and table dashpots, yield with
    unloading curves, friction with and without stick-slip, inactive."""
    from parmec_amd import parmec as P
    rng = np.random.default_rng(seed)
    P.RESET()
    mat = P.MATERIAL(1000.0, 1e6, 0.2)
    pos = rng.uniform(-1.0, 1.0, size=(n, 3)) * np.array([2.0, 2.0, 1.0])
    for i in range(n):
        P.SPHERE(tuple(float(x) for x in pos[i]), 0.1, mat, 1)
        P.VELOCITY(i, tuple(float(x) for x in rng.normal(size=3) * 0.5), tuple(float(x) for x in rng.normal(size=3)))
    P.GRANULAR(0, 0, 1e5, 0.2, 0.3)
    for k in range(3 * n):
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        if k % 7 == 6:
            b = -1
        x1 = tuple(float(x) for x in pos[a] + rng.normal(size=3) * 0.05)
        x2 = tuple(float(x) for x in (pos[b] if b >= 0 else pos[a] + rng.normal(size=3)) + rng.normal(size=3) * 0.05)
        s = sorted(rng.uniform(-0.5, 0.5, size=4))
        f = np.cumsum(rng.uniform(1e3, 1e5, size=4) * np.diff([-0.6] + s)) - 2e3
        curve = [v for pair in zip(s, f) for v in (float(pair[0]), float(pair[1]))]
        kw = {}
        mode = k % 4
        if mode == 1:
            kw["direction"] = tuple(float(x) for x in rng.normal(size=3))
        elif mode == 2:
            kw["direction"] = tuple(float(x) for x in rng.normal(size=3))
            kw["planar"] = "ON"
        if k % 3 == 0:
            kw["dashpot"] = float(rng.uniform(0.0, 0.5))
        elif k % 3 == 1:
            kw["dashpot"] = [-10.0, -50.0, 10.0, 50.0]
        if k % 5 == 2:
            kw["ylim"] = (-float(rng.uniform(1e2, 1e4)), float(rng.uniform(1e2, 1e4)))
            kw["unload"] = [-1.0, -4e4, 0.0, 0.0, 1.0, 4e4]
        if k % 4 == 3:
            kw["friction"] = 0.3
            kw["kskn"] = 0.0 if k % 8 == 3 else 0.5
        if k % 11 == 10:
            kw["inactive"] = True
        geom2 = x2
        if mode == 3 and b < 0:
            geom2 = [x2, tuple(float(x) for x in rng.normal(size=3))]
        P.SPRING(a, x1, b, geom2, curve, **kw)
    m = P._build_model()
    P.RESET()
    return m
