"""The FP32 pre-test of the detection kernel (near32 in pmx_step.hip) must
never reject a pair that the FP64 sphere test accepts.  This restates both
tests in numpy (IEEE binary32/binary64, round to nearest, as the device
arithmetic; contraction is off on the device) and drives them with
adversarial pairs: touching to within 1e-12 of r_a + r_b in random
directions, centres far from the origin O relative to the radii, radius
ratios up to 1e3, and scales from 1e-6 to 1e6.  It also checks that the
pre-test does reject clear misses (it is a filter, not a pass-through)."""
import numpy as np

U = 2.0 ** -24
K1 = np.float32(1.0 + 2.0 ** -20)


def fp64_accepts(a, ra, b, rb):
    """overlap() of pmx_step.hip: d2 < sr^2 (1 + 1e-10) and sr - sqrt(d2) > 0."""
    q = a - b
    d2 = q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2]
    sr = ra + rb
    return (d2 < sr * sr * 1.0000000001) & (sr - np.sqrt(d2) > 0.0)


def near32(a, ra, b, rb, O, B):
    """near32() with the k2 of k_detect: offsets from O narrowed to FP32."""
    k2 = np.where(B > 1e15, np.float32(np.inf),
                  (2.0 ** -21 * B * (1.0 + 1e-6)).astype(np.float32) + np.float32(1e-18))
    pa = (a - O).astype(np.float32)
    pb = (b - O).astype(np.float32)
    d = pa - pb
    d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    thr = (ra.astype(np.float32) + rb.astype(np.float32)) * K1 + k2
    return ~(d2 > thr * thr)


def _case(rng, n, scale, spread, ratio):
    """Block sphere a at offset up to spread*scale from O, partner b touching
    it (distance (r_a + r_b)(1 - eps)) in a random direction."""
    O = rng.uniform(-1e3, 1e3, size=(n, 3)) * scale
    ra = scale * rng.uniform(0.5, 1.5, size=n)
    rb = ra * np.exp(rng.uniform(-np.log(ratio), np.log(ratio), size=n))
    a = O + rng.uniform(-1.0, 1.0, size=(n, 3)) * spread * scale
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    eps = 10.0 ** rng.uniform(-12, -3, size=n)
    b = a + u * ((ra + rb) * (1.0 - eps))[:, None]
    # B: the bound on |c_a - O| per axis that k_detect derives from its box
    B = np.max(np.abs(a - O), axis=1) * (1.0 + 1e-12)
    return a, ra, b, rb, O, B


def test_pretest_keeps_every_accepted_pair():
    rng = np.random.default_rng(7)
    total = 0
    for scale in (1e-6, 1e-3, 1.0, 1e3, 1e6):
        for spread in (0.0, 1.0, 30.0, 1e4):
            for ratio in (1.0, 10.0, 1e3):
                a, ra, b, rb, O, B = _case(rng, 20000, scale, spread, ratio)
                acc = fp64_accepts(a, ra, b, rb)
                keep = near32(a, ra, b, rb, O, B)
                assert acc.sum() > 0
                bad = acc & ~keep
                assert not bad.any(), (scale, spread, ratio, np.flatnonzero(bad)[:5])
                total += int(acc.sum())
    assert total > 500000


def test_pretest_rejects_clear_misses():
    rng = np.random.default_rng(8)
    a, ra, b, rb, O, B = _case(rng, 20000, 1.0, 30.0, 1.0)
    u = (b - a) / np.linalg.norm(b - a, axis=1)[:, None]
    far = a + u * ((ra + rb) * 1.001)[:, None]      # 0.1 % beyond contact
    assert not fp64_accepts(a, ra, far, rb).any()
    assert not near32(a, ra, far, rb, O, B).any()


def test_pretest_switches_off_for_huge_offsets():
    a = np.array([[1e20, 0.0, 0.0]])
    b = np.array([[1e20 + 1.0, 0.0, 0.0]])
    r = np.array([0.75])
    O = np.zeros((1, 3))
    B = np.array([1e20])
    assert fp64_accepts(a, r, b, r).all()
    assert near32(a, r, b, r, O, B).all()
