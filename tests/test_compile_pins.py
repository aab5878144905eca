"""Pins of the compile step's orderings (renumber.cpp:30-143, batches.cpp:21-50) that do not
restate the algorithm: RCM node orders and element orders derived by hand on small graphs where
the George-Liu root, the (degree, index) neighbour order, the component order and self loops
decide the result, and the greedy colouring checked by brute force against its stated rule.

Hand derivations (node ids as in each model; deg = neighbours in the node graph, which joins
every pair of nodes of an element and every equalDOF pair, sorted and unique, a node listed twice
in one element keeping itself as a neighbour):

2x2x1 grid (18 nodes, id i + 3j + 9k): corners deg 7, edge middles 11, centre column 17. Search
start 0; BFS from 0 ends at depth 2 on {2,5,6,7,8,11,14,15,16,17}; smallest (deg, id) there is 2;
BFS from 2 again has depth 2, so the root is 2. Cuthill-McKee from 2: [2 | 11 1 5 10 14 4 13 |
(from 1) 0 9 3 12 | (from 5) 8 17 7 16 | (from 4) 6 15], reversed. Elements (file order e = ei +
2 ej) keyed by their smallest rank: e2 0, e3 2, e0 6, e1 10.

Tied 1x1x3 column (16 nodes, id i + 2j + 4k; side nodes of each level tied to the first): levels
0 and 3 deg 7, levels 1 and 2 deg 11 (the ties join nodes that already share elements). Start 0,
BFS depth 3 ending at {12..15}: root 12; from 12: 13 14 15 (deg 7) before 8 9 10 11 (deg 11), then
4 5 6 7, then 0 1 2 3; reversed.

Two bricks listed Q then P, P on the even ids 0..14 and Q on the odd ids 1..15 (two components,
each K8): components in order of their smallest id, so P first; in each the root is the second
smallest id (the BFS from the smallest ends on all others); CM then lists the rest by id:
P: [2 0 4 6 8 10 12 14] reversed, Q: [3 1 5 ... 15] reversed. Q holds rank 8..15, P rank 0..7.

Three bricks A = 0..7, B = 8..15, C = 16..23 joined by equalDOF 7-8 and 1-23 (deg 8 at 1, 7, 8,
23, else 7): start 0; BFS from 0 ends at depth 3 on [16..22, 9..15] (discovered in that order),
whose smallest (deg, id) is 9 (a "first discovered" rule would take 16); from 9 the depth grows
to 5, ending at 16..22: root 16; from 16 the depth stays 5. CM from 16: 17..22 (deg 7) then 23
(deg 8), 1, 0 2 3 4 5 6 (deg 7) then 7 (deg 8), 8, 9..15; reversed.

Two bricks, the second degenerate (its corners 6 and 7 both node 8; nodes 0..10): node 8 keeps
a self loop, deg 7 against 6 for 9 and 10. Start 9 (the smallest deg-6 id); BFS ends at {0,3,4,7}:
root 0, confirmed. CM from 0: 3 4 7 (deg 7) then 1 2 5 6 (deg 10); from 1: 9 10 (deg 6) then 8
(deg 7); reversed. Without the self loop 8 would come before 9 and 10.
"""
import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from oracle import fxo

UNIT = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], float)


def bricks(conns, nnode, ties=(), offsets=None, xyz=None):
    """A model from explicit connectivity; each brick's corners at a unit cube (shifted apart)
    unless xyz is given."""
    if xyz is None:
        xyz = np.zeros((nnode, 3))
        for b, c in enumerate(conns):
            for a, n in enumerate(c):
                xyz[n] = UNIT[a] + np.array([3.0 * b if offsets is None else offsets[b], 0.0, 0.0])
    conn = np.array(conns, np.int32).reshape(-1, 8)
    return dict(nnode=nnode, xyz=xyz, nelem=len(conn), conn=conn, group=np.zeros(len(conn), np.int32),
                fix=np.zeros((0, 2), np.int32), tie=np.array(ties, np.int32).reshape(-1, 3),
                params=[dict(E=1e7, nu=0.3, rho=1.0)])


def grid_2x2x1():
    xyz, conn, _ = S.grid(2, 2, 1, 1.0, 1.0, 1.0)
    return dict(nnode=len(xyz), xyz=xyz, nelem=len(conn), conn=conn, group=np.zeros(len(conn), np.int32),
                fix=np.zeros((0, 2), np.int32), tie=np.zeros((0, 3), np.int32), params=[dict(E=1e7, nu=0.3)])


def tied_column():
    return S.soil_column(3, 1.0, params=[dict(E=1e7, nu=0.3)])


def two_components():
    P = [0, 2, 4, 6, 8, 10, 12, 14]
    Q = [1, 3, 5, 7, 9, 11, 13, 15]
    return bricks([Q, P], 16)


def three_tied_bricks():
    A, B, C = list(range(8)), list(range(8, 16)), list(range(16, 24))
    return bricks([A, B, C], 24, ties=[(7, 8, 0), (1, 23, 0)])


def degenerate_pair():
    E0 = [0, 1, 2, 3, 4, 5, 6, 7]
    E1 = [1, 9, 10, 2, 5, 8, 8, 6]
    xyz = np.zeros((11, 3))
    xyz[:8] = UNIT
    xyz[9], xyz[10], xyz[8] = (2, 0, 0), (2, 1, 0), (2, 0, 1)
    return bricks([E0, E1], 11, xyz=xyz)


CASES = {
    "grid_2x2x1": (grid_2x2x1, [15, 6, 16, 7, 17, 8, 12, 3, 9, 0, 13, 4, 14, 10, 5, 1, 11, 2], [2, 3, 0, 1]),
    "tied_column_1x1x3": (tied_column, [3, 2, 1, 0, 7, 6, 5, 4, 11, 10, 9, 8, 15, 14, 13, 12], [0, 1, 2]),
    "two_components": (two_components, [14, 12, 10, 8, 6, 4, 0, 2, 15, 13, 11, 9, 7, 5, 1, 3], [1, 0]),
    "three_tied_bricks": (three_tied_bricks,
                          [15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 0, 1, 23, 22, 21, 20, 19, 18, 17, 16],
                          [1, 0, 2]),
    "degenerate_self_loop": (degenerate_pair, [8, 10, 9, 6, 5, 2, 1, 7, 4, 3, 0], [1, 0]),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("side", ["oracle", "product"])
def test_rcm_and_element_order_match_the_hand_derivation(name, side):
    build, node_order, elem_order = CASES[name]
    m = build()
    c = fxo.OracleMesh(m) if side == "oracle" else core.Mesh(m)
    np.testing.assert_array_equal(c.node_order, node_order)
    np.testing.assert_array_equal(c.node_rank[node_order], np.arange(len(node_order)))
    np.testing.assert_array_equal(c.elem_order, elem_order)


def brute_force_colours(eq, count, padded):
    """Each element, in batch order, takes the smallest colour that no earlier element sharing
    one of its equations holds (pairwise check over all earlier elements)."""
    E = eq.reshape(24, padded)[:, :count]
    eqsets = [set(int(q) for q in E[:, e] if q >= 0) for e in range(count)]
    colour = np.full(count, -1, np.int64)
    for e in range(count):
        taken = {colour[f] for f in range(e) if eqsets[f] & eqsets[e]}
        c = 0
        while c in taken:
            c += 1
        colour[e] = c
    return colour


COLOUR_MODELS = {
    "grid_2x2x1": grid_2x2x1,
    "tied_column_1x1x3": tied_column,
    "three_tied_bricks": three_tied_bricks,
    "layered_box_8": lambda: S.config_c2(n=8),
    "tied_column_2x2x6": lambda: S.soil_column(6, 1.0, params=[dict(E=1e7, nu=0.3)], nx=2, ny=2),
}


@pytest.mark.parametrize("name", sorted(COLOUR_MODELS))
@pytest.mark.parametrize("side", ["oracle", "product"])
def test_greedy_colouring_by_brute_force(name, side):
    m = COLOUR_MODELS[name]()
    c = fxo.OracleMesh(m) if side == "oracle" else core.Mesh(m)
    want = brute_force_colours(c.eq, c.nelem, c.padded)
    start, elems = np.asarray(c.color_start), np.asarray(c.color_elements)
    assert start[0] == 0 and start[-1] == c.nelem and len(start) == want.max() + 2
    got = np.full(c.nelem, -1, np.int64)
    for col in range(len(start) - 1):
        members = elems[start[col]:start[col + 1]]
        assert np.all(np.diff(members) > 0), "batch order within a colour"
        got[members] = col
    np.testing.assert_array_equal(got, want)
    # no two elements of one colour share an equation (an element may hold one equation at
    # several DOFs: tied nodes)
    E = c.eq.reshape(24, c.padded)
    for col in range(len(start) - 1):
        q = np.concatenate([np.unique(E[:, e][E[:, e] >= 0]) for e in elems[start[col]:start[col + 1]]])
        assert len(q) == len(np.unique(q))
