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
