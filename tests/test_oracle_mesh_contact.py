"""CPU pins of sphere contact with a meshed particle's triangles in the
oracle (condet.ispc:412-516 with triobs >= 0; forces.ispc:219-262 partner
velocity of particle j, reduced mass 1/(1/m_i + 1/m_j); the slave copy of
383-405): a sphere hitting a free meshed block head-on rebounds with the
restitution of the granular law for the pair's reduced mass, momentum is
conserved, and at every step in contact the block's force is exactly minus
the sphere's (no gravity).  The sphere lands on the top face's diagonal,
the edge its two triangles share: both make the same contact (as in the
reference), so stiffness and damping double and the restitution is that of
zeta * sqrt(2); the force passes through the block's centre of mass, which
does not turn."""
import math

import numpy as np
import pytest

from oracle import orc
from parmec_amd import parmec as P, pmx

HEX = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def block_deck(zeta, v0=-1.0, gap=0.002, x=(0.0, 0.0)):
    """a sphere of radius 0.5 just above a free 2 x 2 x 0.5 meshed block
    (above x), moving down at v0; no gravity, no friction"""
    P.RESET()
    mat = P.MATERIAL(1000.0, 1e6, 0.25)
    nodes = [c for i, j, k in HEX for c in (-1.0 + 2.0 * i, -1.0 + 2.0 * j, -0.5 + 0.5 * k)]
    b = P.MESH(nodes, [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 2)
    s = P.SPHERE((x[0], x[1], 0.5 + gap), 0.5, mat, 1)
    P.GRANULAR(0, 0, 1e7, zeta, 0.0)
    P.VELOCITY(s, (0.0, 0.0, v0), (0.0, 0.0, 0.0))
    return b, s


@pytest.fixture(autouse=True)
def fresh():
    P.RESET()
    yield
    P.RESET()


@pytest.mark.parametrize("zeta", [0.1, 0.3])
def test_sphere_on_free_block_restitution_and_momentum(zeta):
    b, s = block_deck(zeta)
    P._use_engine(orc.Oracle(1))
    m0 = P.state().copy()
    mb, ms = m0.mass[b], m0.mass[s]
    p0 = mb * m0.linear[:, b] + ms * m0.linear[:, s]
    mred = 1.0 / (1.0 / mb + 1.0 / ms)
    wn = math.sqrt(2.0 * 1e7 / mred)
    h = 0.02 * 2.0 / wn
    P.DEM(0.002 / 1.0 + 3.0 * math.pi / wn, h)
    m = P.state()
    assert len(P.contacts()) == 0
    vrel = m.linear[2, s] - m.linear[2, b]
    z = zeta * math.sqrt(2.0)
    e = math.exp(-z * math.pi / math.sqrt(1 - z * z))
    assert abs(vrel / 1.0 - e) < 0.02 * e
    assert np.abs(m.angular[:, b]).max() < 1e-12
    p1 = mb * m.linear[:, b] + ms * m.linear[:, s]
    assert np.abs(p1 - p0).max() < 1e-12 * np.abs(p0).max()
    assert m.linear[2, b] < 0.0                 # the block is pushed down


def test_action_equals_reaction_every_step():
    b, s = block_deck(0.2, gap=-0.001, x=(0.6, -0.1))   # in contact from the start, off centre
    o = orc.Oracle(1)
    P._use_engine(o)
    seen = 0
    for _ in range(20):
        P.DEM(2e-5, 1e-5)
        m = P.state()
        n = len(P.contacts())
        if n:
            seen += 1
            assert np.array_equal(m.force[:, b], -m.force[:, s])
    assert seen == 20
