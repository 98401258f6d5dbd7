"""The oracle's broad phase pinned without the oracle's own code.

The contact-pair set depends on the exact partitioning tree (the reference's
frozen k-d tree misses some overlapping pairs, and the GPU path reproduces
which).  Here the tree is rebuilt from its definition, independently of the
oracle's Karras-style construction (per-node range search with delta(),
partition.ispc:268-394): the binary radix tree of the sorted 64-bit keys
(Morton code << 32 | sorted position), split recursively at the highest
differing bit, a node of at most CUTOFF = 64 points terminal
(partition.h:29), split dimension (prefix length - 2) % 3, split coordinate
the minimum of the right child's points along it (partition.ispc:297-308,
385); centres descend left when c[d] < coord (600-646); a box reaches the
left child when lo[d] <= coord and the right one when hi[d] >= coord
(condet.ispc:104-113).  The oracle's leaf of every centre and the leaves
every box reaches must be exactly these, on inputs with equal codes, equal
coordinates and boxes whose bounds sit exactly on split coordinates; and a
tree built with a plausible misreading (CUTOFF 32, split coordinate = max of
the left child, strict traversal compares, another dimension rule) must
differ from the oracle on them.

Then two hand-built cases derive, on paper, where the reference rule misses
an overlapping pair: a sphere pair whose boxes do not reach each
other's leaves across a two-level split, and a triangle whose box is not
inflated by the sphere radius (condet.ispc:412-426, 510-511).
"""
import numpy as np
import pytest

from parmec_amd import pmx, workloads

CUTOFF = 64


def _oracle():
    from oracle import orc
    return orc.load()


def _spread10(v):
    out = np.zeros_like(v)
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def morton(P):
    """30-bit codes, x most significant, 1024 cells per axis of the extents
    (partition.ispc:69-94); a zero extent maps the axis to cell 0."""
    lo, hi = P.min(axis=0), P.max(axis=0)
    w = hi - lo
    q = np.zeros(P.shape, dtype=np.int64)
    for k in range(3):
        p = (P[:, k] - lo[k]) / w[k] if w[k] > 0 else np.zeros(len(P))
        q[:, k] = np.clip(p * 1024.0, 0.0, 1023.0).astype(np.int64)
    return _spread10(q[:, 0]) * 4 + _spread10(q[:, 1]) * 2 + _spread10(q[:, 2])


def build_tree(P, cutoff=CUTOFF, coord_rule="min_right", dim_rule=lambda p: (p - 2) % 3, reverse_ties=False):
    """Nodes ("node", dim, coord, left, right) / ("leaf", id), root first;
    leaves numbered left-first depth-first."""
    n = len(P)
    code = morton(P)
    if reverse_ties:   # equal codes in descending index order (an unstable sort)
        order = np.lexsort((-np.arange(n), code))
    else:
        order = np.argsort(code, kind="stable")
    keys = [(int(code[order[i]]) << 32) | i for i in range(n)]
    nodes = []
    nleaves = [0]

    def clz64(x):
        return 64 - x.bit_length()

    def build(a, b):
        me = len(nodes)
        nodes.append(None)
        if b - a + 1 <= cutoff:
            nodes[me] = ("leaf", nleaves[0])
            nleaves[0] += 1
            return me
        prefix = clz64(keys[a] ^ keys[b])
        g = max(i for i in range(a, b) if clz64(keys[a] ^ keys[i]) > prefix)
        d = dim_rule(prefix)
        if coord_rule == "min_right":
            coord = min(P[order[k], d] for k in range(g + 1, b + 1))
        else:
            coord = max(P[order[k], d] for k in range(a, g + 1))
        left = build(a, g)
        right = build(g + 1, b)
        nodes[me] = ("node", d, coord, left, right)
        return me

    build(0, n - 1)
    return nodes


def leaf_of(nodes, p):
    i = 0
    while nodes[i][0] == "node":
        _, d, c, left, right = nodes[i]
        i = left if p[d] < c else right
    return nodes[i][1]


def reach(nodes, lo, hi, strict=False):
    out = []

    def go(i):
        nd = nodes[i]
        if nd[0] == "leaf":
            out.append(nd[1])
            return
        _, d, c, left, right = nd
        if (lo[d] < c) if strict else (lo[d] <= c):
            go(left)
        if (hi[d] > c) if strict else (hi[d] >= c):
            go(right)

    go(0)
    return out


def point_sets():
    """Inputs of 65-300 points: uniform, exact duplicates (equal codes and
    coordinates), a sub-cell cluster (equal codes, distinct coordinates), an
    integer lattice (equal coordinates along every axis), a flat set (zero
    z extent), and a mix."""
    rng = np.random.default_rng(11)
    sets = {}
    for n in (65, 100, 200, 300):
        sets[f"uniform{n}"] = rng.uniform(-1.0, 2.0, size=(n, 3))
    base = rng.uniform(0.0, 1.0, size=(12, 3))
    sets["duplicates150"] = base[rng.integers(0, 12, size=150)]
    cl = np.concatenate([0.5 + rng.uniform(0.0, 1e-5, size=(120, 3)), rng.uniform(0.0, 1.0, size=(60, 3))])
    sets["subcell_cluster180"] = cl
    g = np.stack(np.meshgrid(np.arange(5.0), np.arange(5.0), np.arange(8.0), indexing="ij"), axis=-1).reshape(-1, 3)
    sets["lattice200"] = g
    flat = rng.uniform(0.0, 1.0, size=(130, 3))
    flat[:, 2] = 0.25
    sets["flat130"] = flat
    mix = np.concatenate([g[:100] * 0.1, base[rng.integers(0, 12, size=60)], rng.uniform(0.0, 0.5, size=(90, 3))])
    sets["mix250"] = mix
    return sets


def boxes_for(P, seed=0):
    """Per point: a radius box, the point itself, and boxes with one bound
    exactly on the point's coordinates (split coordinates are point
    coordinates, so some of these bounds sit exactly on a split)."""
    rng = np.random.default_rng(seed)
    span = float((P.max(axis=0) - P.min(axis=0)).max()) or 1.0
    out = []
    for p in P:
        r = rng.uniform(0.01, 0.2) * span
        out.append(np.concatenate([p - r, p + r]))
        out.append(np.concatenate([p, p]))
        d = rng.uniform(0.0, 0.1) * span
        out.append(np.concatenate([p, p + d]))
        out.append(np.concatenate([p - d, p]))
    return np.array(out)


def oracle_tree(P, B):
    import ctypes as C
    lib = _oracle()
    n = len(P)
    c = np.ascontiguousarray(P.T)
    leafof = np.full(n, -1, dtype=np.int32)
    ovf = C.c_int(0)
    nl = lib.orc_tree_leaves(n, c.ctypes.data_as(C.POINTER(C.c_double)), leafof.ctypes.data_as(C.POINTER(C.c_int)),
                             C.byref(ovf))
    box = np.ascontiguousarray(B)
    off = np.zeros(len(B) + 1, dtype=np.int64)
    cap = 64 * len(B)
    lv = np.zeros(cap, dtype=np.int32)
    tot = lib.orc_tree_reach(n, c.ctypes.data_as(C.POINTER(C.c_double)), len(B), box.ctypes.data_as(C.POINTER(C.c_double)),
                             off.ctypes.data_as(C.POINTER(C.c_longlong)), lv.ctypes.data_as(C.POINTER(C.c_int)), cap)
    assert tot <= cap
    return nl, leafof, [list(lv[off[b]:off[b + 1]]) for b in range(len(B))]


def definition(P, B, **kw):
    strict = kw.pop("strict", False)
    nodes = build_tree(P, **kw)
    nl = sum(1 for nd in nodes if nd[0] == "leaf")
    leafof = np.array([leaf_of(nodes, p) for p in P], dtype=np.int32)
    return nl, leafof, [reach(nodes, b[:3], b[3:], strict=strict) for b in B]


SETS = point_sets()


@pytest.mark.parametrize("name", sorted(SETS))
def test_oracle_tree_is_the_definition(name):
    P = SETS[name]
    B = boxes_for(P)
    nl_o, leaf_o, reach_o = oracle_tree(P, B)
    nl_d, leaf_d, reach_d = definition(P, B)
    assert nl_o == nl_d and nl_o > 1
    np.testing.assert_array_equal(leaf_o, leaf_d)
    assert reach_o == reach_d


MUTATIONS = {
    "cutoff32": dict(cutoff=32),
    "max_of_left": dict(coord_rule="max_left"),
    "strict_traversal": dict(strict=True),
    "dimension_(p-1)%3": dict(dim_rule=lambda p: (p - 1) % 3),
    "dimension_p%3": dict(dim_rule=lambda p: p % 3),
}


@pytest.mark.parametrize("mutation", sorted(MUTATIONS))
def test_misreadings_differ_from_oracle(mutation):
    """Each plausible misreading of the tree rules, built by the definition
    above, disagrees with the oracle (leaf membership or reached leaves) on
    the inputs above: the comparison pins those rules."""
    differs = []
    for name, P in SETS.items():
        B = boxes_for(P)
        o = oracle_tree(P, B)
        d = definition(P, B, **dict(MUTATIONS[mutation]))
        if o[0] != d[0] or not np.array_equal(o[1], d[1]) or o[2] != d[2]:
            differs.append(name)
    assert differs, f"{mutation} is not distinguished by any input"


# ------------------------------------------- pairs the frozen tree misses

def q10_model(b_x=5.05, tri_y=None, c_y=4.5):
    """256 spheres in x, y in [0, 8] (z in [0, 8]) that the tree splits into
    four leaves of 64 (x at the most significant Morton bit, then y):

      LL (x < 4, y < 4): A = (3.95, 3.0, 4), r = 1, and 63 small fillers;
      LH (x < 4, y >= 4.5): 64 fillers;  its y split is at min y = 4.5;
      RL (x >= 4, y < 4): 64 fillers, one at x = 4.0 exactly, so the root
          splits x at 4.0;
      RH (x >= 4, y >= 4.1): B = (b_x, 4.1, 4), r = 1, so the right
          subtree's y split is at 4.1; C = (7, c_y, 4), r = 0.5; 62 fillers.

    Fillers have r = 0.01 and touch nothing.  With b_x = 5.05, A and B
    overlap (|AB| = 1.556 < 2) but A's box [2.95, 4.95] x [2, 4] reaches LL
    and RL only (4 < 4.1 at the right y split) and B's box [4.05, 6.05] x
    [3.1, 5.1] reaches RL and RH only (4.05 > 4.0 at the root): the pair is
    never tested.  With b_x = 5.0, B's lower x bound equals the root's split
    coordinate, 4.0 <= 4.0 takes B left, and the pair is found.  A triangle
    lying in the plane y = tri_y under C (x in [6.6, 7.6], z in [3, 5]): its
    box is not widened by C's radius, so at tri_y = 4.05 it reaches RL only
    (4.05 < 4.1) and C (distance 0.45 < 0.5 from it) is never tested; at
    tri_y = 4.1 it reaches RH too and the contact is found."""
    rng = np.random.default_rng(3)
    pts, rad = [], []

    def grid(xs, ys):
        for x in xs:
            for y in ys:
                pts.append((x, y, rng.uniform(0.0, 8.0)))
                rad.append(0.01)

    pts.append((3.95, 3.0, 4.0)); rad.append(1.0)          # A: sphere 0
    pts.append((b_x, 4.1, 4.0)); rad.append(1.0)           # B: sphere 1
    pts.append((7.0, c_y, 4.0)); rad.append(0.5)           # C: sphere 2
    grid(np.linspace(0.0, 2.5, 7), np.linspace(0.0, 3.8, 9))              # LL, 63
    grid(np.linspace(0.0, 2.5, 8), np.linspace(4.5, 8.0, 8))              # LH, 64
    pts.append((4.0, 0.5, 2.0)); rad.append(0.01)                        # RL: the root's split
    grid(np.linspace(6.5, 8.0, 7), np.linspace(0.0, 3.8, 9))              # RL, 63 more
    xs, ys = np.linspace(6.5, 8.0, 7), np.linspace(6.0, 8.0, 9)
    grid(xs, ys[:-1])                                                     # RH, 56
    for x in xs[1:]:                                                      # RH, 6 more
        pts.append((x, 8.0, rng.uniform(0.0, 8.0))); rad.append(0.01)
    P = np.array(pts)
    P[3, 2] = 0.0   # z extents exactly [0, 8]
    P[4, 2] = 8.0
    tris = [] if tri_y is None else [(6.6, tri_y, 3.0, 7.6, tri_y, 3.0, 7.1, tri_y, 5.0)]
    n = len(P)
    m = pmx.Model(parnum=n, ellnum=n, trinum=len(tris), pairnum=1)
    workloads._spheres(m, P, np.array(rad), 1.0)
    if tris:
        m.tri[:] = np.array(tris).T
        m.tricol[:] = 2
    workloads._granular(m, 1e3, 0.0, 0.0)
    return m


def brute_pairs(m):
    c, r = m.center[:3].T, m.radii[0]
    out = set()
    for i in range(len(r)):
        d = np.sqrt(((c - c[i]) ** 2).sum(axis=1))
        for j in np.nonzero(r[i] + r - d > 0.0)[0]:
            if j > i:
                out.add((i, int(j)))
    return out


def q10_contacts(engine, m):
    mm = m.copy()
    engine.upload(mm)
    engine.dem(0.5e-4, 1e-4, pmx.forcing())   # one step from rest
    cs = engine.contacts()
    return {(int(a), int(b)) for a, b in zip(cs["master"], cs["slave"])}


def test_q10_layout_is_as_derived():
    """The four leaves and split coordinates the derivation assumes."""
    m = q10_model()
    P = m.center[:3].T.copy()
    nodes = build_tree(P)
    assert nodes[0][0] == "node" and nodes[0][1] == 0 and nodes[0][2] == 4.0
    left, right = nodes[nodes[0][3]], nodes[nodes[0][4]]
    assert left[:3] == ("node", 1, 4.5) and right[:3] == ("node", 1, 4.1)
    nl, leafof, _ = oracle_tree(P, np.zeros((0, 6)))
    assert nl == 4
    assert list(np.bincount(leafof)) == [64, 64, 64, 64]
    assert leafof[0] == 0 and leafof[1] == 3 and leafof[2] == 3


@pytest.mark.parametrize("b_x,found", [(5.05, False), (5.0, True)])
def test_q10_sphere_pair(b_x, found):
    from oracle import orc
    m = q10_model(b_x=b_x)
    assert brute_pairs(m) == {(0, 1)}
    o = orc.Oracle(0)
    got = q10_contacts(o, m)
    o.close()
    assert got == ({(0, 1)} if found else set())


@pytest.mark.parametrize("tri_y,c_y,found", [(4.05, 4.5, False), (4.1, 4.55, True)])
def test_q10_triangle_box_not_inflated(tri_y, c_y, found):
    from oracle import orc
    m = q10_model(tri_y=tri_y, c_y=c_y)
    depth = 0.5 - (c_y - tri_y)
    assert depth > 0.0
    o = orc.Oracle(0)
    got = q10_contacts(o, m)
    o.close()
    assert (0, 1) not in got   # b_x = 5.05: the sphere pair is missed
    assert got == ({(2, -1)} if found else set())
