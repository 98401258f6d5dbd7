"""MESH in the deck API (input.cpp:840-1156, mesh.cpp:315-521): surface
triangles of a volume mesh and its mass characteristics, against closed
forms; particle numbering shared with SPHERE; argument errors."""
import math

import numpy as np
import pytest

from parmec_amd import parmec as P

HEX = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


@pytest.fixture(autouse=True)
def fresh():
    P.RESET()
    yield
    P.RESET()


def flat(points):
    return [float(c) for p in points for c in p]


def cuboid_nodes(x0, size):
    return [(x0[0] + i * size[0], x0[1] + j * size[1], x0[2] + k * size[2]) for i, j, k in HEX]


def outward(m, part, center):
    T = m.tri.T.reshape(-1, 3, 3)[m.triobs == part]
    for t in T:
        n = np.cross(t[1] - t[0], t[2] - t[1])
        assert n @ (t.mean(axis=0) - center) > 0.0
    return len(T)


def test_cuboid_mass_properties_and_surface():
    rho, size, x0 = 1800.0, (0.2, 0.2, 0.4), (1.0, 2.0, 3.0)
    mat = P.MATERIAL(rho, 1e9, 0.25)
    p = P.MESH(flat(cuboid_nodes(x0, size)), [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 4)
    m = P._build_model()
    a, b, c = size
    M = rho * a * b * c
    assert p == 0 and m.parnum == 1 and m.ellnum == 0
    assert abs(m.mass[0] / M - 1.0) < 1e-14
    np.testing.assert_allclose(m.position[:3, 0], [x0[0] + a / 2, x0[1] + b / 2, x0[2] + c / 2], rtol=1e-15)
    J = m.inertia[:, 0].reshape(3, 3)
    np.testing.assert_allclose(np.diag(J), [M * (b * b + c * c) / 12, M * (a * a + c * c) / 12, M * (a * a + b * b) / 12],
                               rtol=1e-12)
    assert np.abs(J - np.diag(np.diag(J))).max() < 1e-12 * np.trace(J)
    assert m.trinum == 12 and outward(m, 0, m.position[:3, 0]) == 12
    assert set(m.tricol) == {4}
    assert m.nodnum == 8 and m.facnum == 12 and set(m.nodpart) == {0}
    # faces point at the triangles holding their nodes
    for f in range(m.facnum):
        t = m.factri[f]
        np.testing.assert_array_equal(m.nodes[:3, m.facnod[:, f]].T.reshape(-1), m.tri[[0, 1, 2, 3, 4, 5, 6, 7, 8], t])


def test_two_hexes_share_a_face():
    """The shared face is inner: 10 quadrilaterals (20 triangles) remain;
    the mass is the sum; the face colour list overrides the global colour."""
    nodes = cuboid_nodes((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)) + [(2.0, 0.0, 0.0), (2.0, 1.0, 0.0), (2.0, 0.0, 1.0),
                                                                (2.0, 1.0, 1.0)]
    mat = P.MATERIAL(1000.0, 1e9, 0.25)
    els = [8, 0, 1, 2, 3, 4, 5, 6, 7, mat, 8, 1, 8, 9, 2, 5, 10, 11, 6, mat]
    colors = [3, 4, 8, 9, 11, 10, 7]          # the far x = 2 face gets colour 7
    P.MESH(flat(nodes), els, mat, colors)
    m = P._build_model()
    assert abs(m.mass[0] - 2000.0) < 1e-11
    np.testing.assert_allclose(m.position[:3, 0], [1.0, 0.5, 0.5], rtol=1e-15, atol=1e-15)
    assert m.trinum == 20 and outward(m, 0, m.position[:3, 0]) == 20
    assert sorted(m.tricol.tolist()) == [3] * 18 + [7] * 2
    far = m.tri.reshape(3, 3, -1)[:, 0, m.tricol == 7]
    assert np.all(far == 2.0)


def test_tetrahedron_pyramid_wedge():
    """Unit corner tetrahedron: V = 1/6, centroid 1/4, inertia about the
    centroid 1/80 on the diagonal and 1/480 off it; unit pyramid V = 1/3,
    centroid height 1/4; unit triangular prism V = 1/2, centroid (1/3, 1/3, 1/2)."""
    mat = P.MATERIAL(1.0, 1e9, 0.25)
    tet = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    P.MESH(flat(tet), [4, 0, 1, 2, 3, mat], mat, 1)
    pyr = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0.5, 0.5, 1.0)]
    P.MESH(flat(pyr), [5, 0, 1, 2, 3, 4, mat], mat, 1)
    wed = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)]
    P.MESH(flat(wed), [6, 0, 1, 2, 3, 4, 5, mat], mat, 1)
    m = P._build_model()
    np.testing.assert_allclose(m.mass, [1 / 6, 1 / 3, 1 / 2], rtol=1e-14)
    np.testing.assert_allclose(m.position[:3, 0], [0.25, 0.25, 0.25], rtol=1e-14)
    J = m.inertia[:, 0].reshape(3, 3)
    np.testing.assert_allclose(np.diag(J), [1 / 80] * 3, rtol=1e-12)
    np.testing.assert_allclose(J[0, 1], 1 / 480, rtol=1e-12)
    np.testing.assert_allclose(m.position[:3, 1], [0.5, 0.5, 0.25], rtol=1e-14)
    np.testing.assert_allclose(m.position[:3, 2], [1 / 3, 1 / 3, 0.5], rtol=1e-14)
    assert outward(m, 0, m.position[:3, 0]) == 4
    assert outward(m, 1, m.position[:3, 1]) == 6      # one quad, four triangles
    assert outward(m, 2, m.position[:3, 2]) == 8      # two triangles, three quads


def test_numbering_shared_with_spheres():
    mat = P.MATERIAL(1000.0, 1e9, 0.25)
    a = P.SPHERE((5.0, 0.0, 0.0), 0.1, mat, 1)
    b = P.MESH(flat(cuboid_nodes((0, 0, 0), (1, 1, 1))), [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 2)
    c = P.SPHERE((7.0, 0.0, 0.0), 0.1, mat, 1)
    P.OBSTACLE([(0, 0, -1, 1, 0, -1, 0, 1, -1)], 3)
    assert (a, b, c) == (0, 1, 2)
    m = P._build_model()
    assert m.parnum == 3 and m.ellnum == 2
    np.testing.assert_array_equal(m.part, [0, 2])
    np.testing.assert_array_equal(m.triobs, [1] * 12 + [-1])
    assert m.mass[1] == pytest.approx(1000.0) and m.position[0, 2] == 7.0


def test_mesh_argument_errors():
    mat = P.MATERIAL(1000.0, 1e9, 0.25)
    nodes = flat(cuboid_nodes((0, 0, 0), (1, 1, 1)))
    with pytest.raises(ValueError, match="4, 5, 6, or 8"):
        P.MESH(nodes, [7, 0, 1, 2, 3, 4, 5, 6, mat], mat, 1)
    with pytest.raises(ValueError, match="incomplete"):
        P.MESH(nodes, [8, 0, 1, 2, 3, 4, 5, 6, 7], mat, 1)
    with pytest.raises(ValueError, match="outside of range"):
        P.MESH(nodes, [8, 0, 1, 2, 3, 4, 5, 6, 9, mat], mat, 1)
    with pytest.raises(ValueError, match="repeats"):
        P.MESH(nodes, [8, 0, 1, 2, 3, 4, 5, 6, 6, mat], mat, 1)
    with pytest.raises(ValueError, match="multiple of 3"):
        P.MESH(nodes[:-1], [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 1)
    with pytest.raises(ValueError, match="3 or 4 nodes"):
        P.MESH(nodes, [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, [1, 5, 0, 1, 2, 3, 4, 2])
    with pytest.raises(TypeError):
        P.MESH(tuple(nodes), [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 1)


def deck_lattice(nx, ny, nz, size=(0.2, 0.2, 0.4), gap=0.01, density=1800.0, seed=0, stiffness=(1e5, 1e7),
                 curve_step=0.005, ratio=0.05):
    """The spring lattice built through the deck: one MESH hexahedron per
    body, a follower SPRING between facing face centres of neighbours (a
    seeded 5-point curve through 0, critical damping ratio), and the bottom
    layer driven by PRESCRIBE ('av': a sinusoidal base acceleration)."""
    rng = np.random.default_rng(seed)
    a, b, c = size
    mat = P.MATERIAL(density, 1e9, 0.25)
    ids = {}
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                x0 = (i * (a + gap), j * (b + gap), k * (c + gap) + gap)
                ids[i, j, k] = P.MESH(flat(cuboid_nodes(x0, size)), [8, 0, 1, 2, 3, 4, 5, 6, 7, mat], mat, 1)
    h = curve_step
    for (i, j, k), p in ids.items():
        cen = np.array([i * (a + gap) + a / 2, j * (b + gap) + b / 2, k * (c + gap) + c / 2 + gap])
        for d, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            q = ids.get((i + di, j + dj, k + dk))
            if q is None:
                continue
            e = np.zeros(3)
            e[d] = 1.0
            half = 0.5 * np.array(size)[d]
            x1 = tuple(cen + half * e)
            x2 = tuple(cen + (half + gap) * e)
            s = rng.uniform(stiffness[0], stiffness[1], size=4)
            curve = [-2 * h, -(s[1] + s[0]) * h, -h, -s[1] * h, 0.0, 0.0, h, s[2] * h, 2 * h, (s[2] + s[3]) * h]
            P.SPRING(p, x1, q, x2, curve, dashpot=ratio)
    base = [p for (i, j, k), p in ids.items() if k == 0]
    acc = lambda t: (9.81 * math.sin(2.0 * math.pi * 5.0 * t), 0.0, 0.0)
    for p in base:
        P.PRESCRIBE(p, linear=acc, angular=lambda t: (0.0, 0.0, 0.0), kind="av")
    return ids
