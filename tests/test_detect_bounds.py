"""The two shortcuts of the detection kernel (parmec_amd/csrc/pmx_step.hip),
checked on the CPU with the same IEEE round-to-nearest arithmetic:

- f32_below / f32_above: FP32 bounds of a double, stepped out from the
  rounded value by |f| 2^-23 and 2^-120; the reach masks built from them
  must hold every sphere that the FP64 reach test (condet.ispc:104-113)
  accepts, so f32_below(v) <= v <= f32_above(v) for every double v;
- overlap: outside the band sr^2 (1 -/+ 1e-10) the verdict of
  r_a + r_b - sqrt(d2) > 0 (condet.ispc:128-144) is taken without the sqrt.

The expressions are restated below in numpy (float32 / float64 operations,
each correctly rounded, no fused operations, as the kernel is compiled).
"""
import numpy as np

F32_MAX = np.float32(np.finfo(np.float32).max)


def f32_below(v):
    f = v.astype(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        x = f - np.abs(f) * np.float32(2.0 ** -23) - np.float32(2.0 ** -120)
    return np.fmin(x, F32_MAX)


def f32_above(v):
    f = v.astype(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        x = f + np.abs(f) * np.float32(2.0 ** -23) + np.float32(2.0 ** -120)
    return np.fmax(x, -F32_MAX)


def _doubles(rng, n):
    """values across the whole double range, with binade edges, values just
    off them, float rounding midpoints, subnormals and both overflow sides"""
    e = rng.integers(-1074, 1024, n)
    m = rng.random(n) + 1.0
    v = np.ldexp(m, e) * rng.choice([-1.0, 1.0], n)
    edges = np.ldexp(1.0, rng.integers(-160, 140, n)) * rng.choice([-1.0, 1.0], n)
    near = edges * (1.0 + rng.integers(-4, 5, n) * 2.0 ** -52)
    f = rng.standard_normal(n).astype(np.float32).astype(np.float64)
    mid = f + np.ldexp(np.abs(f), -25) * rng.choice([-1.0, 1.0], n)   # half way between floats
    special = np.array([0.0, -0.0, 5e-324, -5e-324, 1e-40, -1e-40, 2.0 ** -126, -(2.0 ** -126),
                        float(F32_MAX), -float(F32_MAX), 3.5e38, -3.5e38, 1e300, -1e300,
                        np.finfo(np.float64).max, -np.finfo(np.float64).max, np.inf, -np.inf])
    return np.concatenate([v, edges, near, f, mid, special])


def test_outward_float_bounds_bracket_every_double():
    rng = np.random.default_rng(7)
    v = _doubles(rng, 400000)
    lo, hi = f32_below(v), f32_above(v)
    assert not np.isnan(lo).any() and not np.isnan(hi).any()
    assert (lo.astype(np.float64) <= v).all()
    assert (hi.astype(np.float64) >= v).all()
    # and the step out stays small: a few float ulps for values in the float range
    ok = np.abs(v) < 1e37
    ok &= np.abs(v) > 1e-25
    rel = (hi[ok].astype(np.float64) - lo[ok].astype(np.float64)) / np.abs(v[ok])
    assert rel.max() < 2.0 ** -20


def test_overlap_shortcuts_match_the_sqrt():
    """pairs drawn on and around the contact distance: the three-way test
    agrees with sr - sqrt(d2) > 0 everywhere"""
    rng = np.random.default_rng(11)
    n = 400000
    sr = np.ldexp(rng.random(n) + 1.0, rng.integers(-40, 40, n))
    s2 = sr * sr
    # d2 spread over the band edges and inside the band, to a few ulps
    rel = np.concatenate([rng.uniform(-3e-10, 3e-10, n // 2),
                          rng.choice([-1e-10, 1e-10], n // 2) * (1.0 + rng.uniform(-1e-6, 1e-6, n // 2))])
    d2 = s2 * (1.0 + rel)
    d2[: 2000] = np.nextafter(s2[: 2000], np.inf)
    d2[2000: 4000] = np.nextafter(s2[2000: 4000], 0.0)
    d2[4000: 6000] = s2[4000: 6000]
    truth = sr - np.sqrt(d2) > 0.0
    far = d2 >= s2 * 1.0000000001
    near = (d2 < s2 * 0.9999999999) & (s2 <= np.finfo(np.float64).max)
    assert not (far & near).any()
    assert not truth[far].any()
    assert truth[near].all()
    # the band really is narrow: most random pairs take a shortcut
    assert (far | near).mean() > 0.3
